"""Identifiers of the C-ABI (include/vectorizedadjoint_amd.h), mirrored for Python."""

# va_system
SYS_VAN_DER_POL = 1
SYS_OSCILLATOR = 2
SYS_OSCILLATOR_FIXED_K = 3
SYS_LOTKA_VOLTERRA = 4
SYS_LOTKA_VOLTERRA_SHARED = 5
SYS_USER = 100  # a system of the caller's, with its own kernel table

# va_tableau
TAB_EULER = 1
TAB_RK4 = 2
TAB_CASH_KARP54 = 3
TAB_DOPRI5 = 4
TAB_FEHLBERG78 = 5

# va_cost
COST_SUM = 1
COST_SQUARED_NORM = 2
COST_LINEAR = 3

# per-instance status
LANE_OK = 0
LANE_MAX_FAIL = 1
LANE_NONFINITE = 2
LANE_MAX_STEPS = 3

# va_mode
MODE_PERSISTENT = 0
MODE_PLAIN = 1

# va_problem.flags
FLAG_GRAPH = 1
FLAG_COLD_ORDER = 2  # every forward takes instances in index order (no schedule reuse)
FLAG_CHECKPOINT_2LEVEL = 4  # two levels of checkpoints even where one fits the budget

# va_status
STATUS_NAMES = {
    0: "VA_OK",
    1: "VA_ERR_INVALID",
    2: "VA_ERR_UNSUPPORTED",
    3: "VA_ERR_HIP",
    4: "VA_ERR_OUT_OF_MEMORY",
    5: "VA_ERR_STATE",
    6: "VA_ERR_LANE_FAILURE",
}
