// Host-side helpers shared by the forward and adjoint translation units of libgsmi.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

#include "../../include/gsmi.h"
#include "device_math.h"
#include "device_program.h"

namespace gsmi {

// Thread-local message returned by gsmi_last_error().
void set_error(const std::string& msg);

// GSMI_OK, or GSMI_E_HIP with the HIP error text recorded.
int hip_check(hipError_t e, const char* what);

// N is 2^k or 3*2^k and the two in-LDS transforms of a workgroup fit `lds` bytes (padded: lines
// in the padded layout of lds_line).
bool fft_size_ok(int N, size_t elem, size_t lds = 65536, bool padded = false);

// LDS bytes one workgroup may allocate on the current device (160 KiB on gfx950; 64 KiB if the
// query fails).  Launches above 64 KiB set the kernel's dynamic-LDS attribute first.
size_t lds_limit();

// hipFuncSetAttribute(MaxDynamicSharedMemorySize) for launches above 64 KiB.
int allow_lds(const void* kernel, size_t bytes);

// Device table exp(+2 pi i m / N), m = 0..N-1, cached per device and N (nullptr on failure).
template <typename T>
const cplx<T>* twiddles(int N);

// The two large twiddle tables of the work-list kernels ({nullptr, nullptr} on failure).
template <typename T>
inline TwiddleSet<T> twiddle_set()
{
    return TwiddleSet<T>{twiddles<T>(kTwPow2), twiddles<T>(kTwPow3)};
}

// The separable axis-factor prepass of every stamp (X(kx), Y(ky) at each stamp's axis_offset).
template <typename T>
int launch_axis_factors(const gsmi_stamp* stamps, int nstamps, const gsmi_op* ops,
                        const gsmi_leaf* leaves, void* scratch, void* stream);

}  // namespace gsmi
