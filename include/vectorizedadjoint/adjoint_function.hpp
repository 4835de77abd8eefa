// adjoint_function.hpp -- batched, autograd-shaped use of the device path.
//
// No reference counterpart: the reference's Driver solves one instance per
// call (lib/include/Driver.hpp:15-78).  AdjointFunction keeps one library
// handle for a problem over B instances and exposes the two halves of a
// differentiable function y(T) = Phi(y0, p):
//   forward(y, p)            y: [N][B] device array, y0 in / y(T) out
//   backward(p, gy, gp, ...) given dJ/dy(T) (gy) and the direct dJ/dp (gp,
//                            may be null), returns dJ/dy0 and dJ/dp
// The backward differentiates the trajectory of the LAST forward on this
// object; a forward in between replaces it.  Arrays are device pointers in
// the C ABI's structure-of-arrays layout; calls are ordered on `stream`.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "../vectorizedadjoint_amd.h"

namespace vectorizedadjoint {

class AdjointFunction
{
  public:
    explicit AdjointFunction(const va_problem &pb) : m_pb(pb)
    {
        check(va_create(&m_pb, &m_h), "va_create");
    }
    AdjointFunction(const AdjointFunction &) = delete;
    AdjointFunction &operator=(const AdjointFunction &) = delete;
    ~AdjointFunction()
    {
        if (m_h) va_destroy(m_h);
    }

    const va_problem &problem() const { return m_pb; }
    va_handle *handle() const { return m_h; }
    uint64_t generation() const { return m_generation; }

    void forward(double *y, const double *p, int32_t *n_accepted = nullptr,
                 int32_t *n_rejected = nullptr, int32_t *lane_status = nullptr,
                 hipStream_t stream = nullptr)
    {
        check(va_forward(m_h, y, p, n_accepted, n_rejected, lane_status, stream), "va_forward");
        ++m_generation;
    }

    // dJdy0 [N][B] = (d y(T)/d y0)^T gy;  dJdp [P][B] = gp + (d y(T)/d p)^T gy.
    // dJdy0 may alias gy; dJdp may alias gp.
    void backward(const double *p, const double *gy, const double *gp, double *dJdy0,
                  double *dJdp, hipStream_t stream = nullptr)
    {
        if (!m_generation) throw std::runtime_error("AdjointFunction: backward before forward");
        const size_t nb = (size_t)m_pb.n_state * (size_t)m_pb.batch;
        const size_t pbk = (size_t)m_pb.n_param * (size_t)m_pb.batch;
        if (dJdy0 != gy)
            hip_check(hipMemcpyAsync(dJdy0, gy, nb * sizeof(double), hipMemcpyDeviceToDevice,
                                     stream));
        if (!gp)
            hip_check(hipMemsetAsync(dJdp, 0, pbk * sizeof(double), stream));
        else if (dJdp != gp)
            hip_check(hipMemcpyAsync(dJdp, gp, pbk * sizeof(double), hipMemcpyDeviceToDevice,
                                     stream));
        check(va_adjoint(m_h, p, 1, dJdy0, dJdp, stream), "va_adjoint");
    }

  private:
    static void check(va_status s, const char *what)
    {
        if (s != VA_OK) throw std::runtime_error(std::string(what) + ": " + va_last_error());
    }
    static void hip_check(hipError_t e)
    {
        if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
    }

    va_problem m_pb;
    va_handle *m_h = nullptr;
    uint64_t m_generation = 0;
};

}  // namespace vectorizedadjoint
