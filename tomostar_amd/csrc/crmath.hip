// Evaluation of the correctly rounded trigonometry (crmath.hpp) over arrays, on the host or on the
// device: the functions GeoMath's positions are made of (GeoMath.cs:49-80), exposed so their results
// can be checked against an independent correctly rounded library.
#include <vector>

#include "common.hpp"
#include "crmath.hpp"

namespace ts {

__host__ __device__ inline double crmath_one(int func, double a, double b)
{
    switch (func) {
    case 0: return cr::sin(a);
    case 1: return cr::cos(a);
    case 2: return cr::asin(a);
    default: return cr::atan2(a, b);
    }
}

__global__ void crmath_kernel(int func, int n, const double *__restrict__ a, const double *__restrict__ b,
                              double *__restrict__ out)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
        out[q] = crmath_one(func, a[q], b ? b[q] : 0.0);
}

}  // namespace ts

using namespace ts;

extern "C" int ts_crmath_eval(ts_context *ctx, int func, int n, const double *a, const double *b, double *out)
{
    if (func < 0 || func > 3 || n < 0 || (n > 0 && (!a || !out)) || (func == 3 && n > 0 && !b)) return TS_EINVAL;
    if (!ctx) {
        for (int q = 0; q < n; q++) out[q] = crmath_one(func, a[q], b ? b[q] : 0.0);
        return TS_OK;
    }
    return guarded(nullptr, 0, [&]() -> int {
        if (n == 0) return TS_OK;
        TS_HIP(hipSetDevice(ctx->device));
        std::lock_guard<std::mutex> lk(ctx->gate);
        hipStream_t st = ctx->stream;
        dbuf<double> da(n), db(func == 3 ? n : 0), dout(n);
        TS_HIP(hipMemcpyAsync(da.p, a, (size_t)n * 8, hipMemcpyHostToDevice, st));
        if (func == 3) TS_HIP(hipMemcpyAsync(db.p, b, (size_t)n * 8, hipMemcpyHostToDevice, st));
        const unsigned blocks = (unsigned)std::min<int64_t>(((int64_t)n + 255) / 256, 8192);
        crmath_kernel<<<blocks, 256, 0, st>>>(func, n, da.p, func == 3 ? db.p : nullptr, dout.p);
        TS_HIP(hipGetLastError());
        TS_HIP(hipMemcpyAsync(out, dout.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
        TS_HIP(hipStreamSynchronize(st));
        return TS_OK;
    });
}
