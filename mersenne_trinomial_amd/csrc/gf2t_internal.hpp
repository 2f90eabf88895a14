// gf2t_internal.hpp -- shared device arithmetic and library plumbing for
// the gfx950 GF(2) trinomial kernels.  Everything here is carry-less
// integer work on 64-wide waves: XOR, shift, and 32x32 integer
// multiplies for the carry-less products (CDNA has no clmul instruction).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>

typedef uint32_t u32;
typedef uint64_t u64;

#define GF2T_DEV __device__ __forceinline__

// ---------------------------------------------------------------- errors
namespace gf2t {
void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);
}  // namespace gf2t

#define GF2T_HIP(call)                                                        \
    do {                                                                      \
        hipError_t e__ = (call);                                              \
        if (e__ != hipSuccess)                                                \
            return gf2t::hip_fail(e__, #call, __FILE__, __LINE__);            \
    } while (0)

#define GF2T_CHECK_LAUNCH(what)                                               \
    do {                                                                      \
        hipError_t e__ = hipGetLastError();                                   \
        if (e__ != hipSuccess)                                                \
            return gf2t::hip_fail(e__, what, __FILE__, __LINE__);             \
    } while (0)

#include "gf2t_arith.hpp"

static inline size_t words_of(u64 bits) { return (size_t)((bits + 63) >> 6); }
