// force.hip -- free-vortex force of bound vortices on the free particle field,
// the pairwise post-processing pass of FLOWUnsteady's _Ffv_direct
// (src/FLOWUnsteady_processing_force.jl:860-929).
//
// Few bound vortices (sources, FLOWUnsteady's static/bound particles) against
// every free particle (targets): the free field is split into segments so the
// grid fills the device whatever the number of bound vortices; each workgroup
// reduces its segment for one bound vortex in a fixed tree order and a second
// kernel adds the segments in order, so results are bitwise repeatable.  FP64.
#include <algorithm>

#include "device_math.h"
#include "internal.h"

namespace vpmx {

namespace {

constexpr int FB = 256;

// bnd: 8 doubles per bound vortex (x, y, z, Gx, Gy, Gz, 1/sigma, 0).
template <int K>
__global__ __launch_bounds__(FB) void ffv_partial_kernel(const double *__restrict__ bnd,
                                                         const double *__restrict__ X,
                                                         const double *__restrict__ G,
                                                         int64_t nf, int64_t cap, int64_t seg,
                                                         int nseg, double *__restrict__ part)
{
    __shared__ double red[6][FB];
    const int b = blockIdx.x / nseg, sg = blockIdx.x - b * nseg;
    const double *B = bnd + 8 * b;
    const double xb = B[0], yb = B[1], zb = B[2], g1 = B[3], g2 = B[4], g3 = B[5], isg = B[6];
    double m0 = 0, m1 = 0, m2 = 0, m3 = 0, m4 = 0, m5 = 0;
    const int64_t f0 = (int64_t)sg * seg, f1 = (f0 + seg < nf) ? f0 + seg : nf;
    for (int64_t f = f0 + threadIdx.x; f < f1; f += FB) {
        const double d1 = X[f] - xb, d2 = X[cap + f] - yb, d3 = X[2 * cap + f] - zb;
        const double r2 = d1 * d1 + d2 * d2 + d3 * d3;
        if (!(r2 > 0.0)) continue;   // the reference skips r == 0
        const double r = sqrt(r2);
        double g, dg;
        g_dg_f64<K>(r * isg, g, dg);
        const double k = g * INV4PI / (r2 * r);
        const double h1 = G[f], h2 = G[cap + f], h3 = G[2 * cap + f];
        // U_b(x_f) = -k dX x Gamma_b;  Ffv += U_b x Gamma_f
        const double u1 = -k * (d2 * g3 - d3 * g2);
        const double u2 = -k * (d3 * g1 - d1 * g3);
        const double u3 = -k * (d1 * g2 - d2 * g1);
        m0 += u2 * h3 - u3 * h2;
        m1 += u3 * h1 - u1 * h3;
        m2 += u1 * h2 - u2 * h1;
        // Fbxf += k dX x (Gamma_b x Gamma_f)
        const double c1 = g2 * h3 - g3 * h2, c2 = g3 * h1 - g1 * h3, c3 = g1 * h2 - g2 * h1;
        m3 += k * (d2 * c3 - d3 * c2);
        m4 += k * (d3 * c1 - d1 * c3);
        m5 += k * (d1 * c2 - d2 * c1);
    }
    red[0][threadIdx.x] = m0; red[1][threadIdx.x] = m1; red[2][threadIdx.x] = m2;
    red[3][threadIdx.x] = m3; red[4][threadIdx.x] = m4; red[5][threadIdx.x] = m5;
    __syncthreads();
    for (int w = FB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * (int64_t)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(FB) void ffv_reduce_kernel(const double *__restrict__ part, int nb,
                                                        int nseg, double *__restrict__ out)
{
    const int i = blockIdx.x * FB + threadIdx.x;   // b * 6 + k
    if (i >= 6 * nb) return;
    const int b = i / 6, k = i - 6 * b;
    double v = 0.0;
    for (int s = 0; s < nseg; ++s) v += part[6 * ((int64_t)b * nseg + s) + k];
    out[i] = v;
}

}  // namespace

int ffv_segments(int64_t nb, int64_t nf)
{
    const int64_t maxseg = std::max<int64_t>(1, (nf + FB - 1) / FB);
    const int64_t want = std::max<int64_t>(1, (2048 + nb - 1) / std::max<int64_t>(nb, 1));
    return (int)std::min(maxseg, want);
}

hipError_t launch_ffv(hipStream_t st, int kernel, const double *bnd, int nb, const DeviceField &f,
                      int nseg, double *part, double *out)
{
    if (nb == 0) return hipSuccess;
    const int64_t seg = (f.np + nseg - 1) / nseg;
    const unsigned g = (unsigned)((int64_t)nb * nseg);
    switch (kernel) {
    case SINGULAR:
        ffv_partial_kernel<SINGULAR><<<g, FB, 0, st>>>(bnd, f.X, f.G, f.np, f.cap, seg, nseg, part);
        break;
    case GAUSSIAN:
        ffv_partial_kernel<GAUSSIAN><<<g, FB, 0, st>>>(bnd, f.X, f.G, f.np, f.cap, seg, nseg, part);
        break;
    case GAUSSIANERF:
        ffv_partial_kernel<GAUSSIANERF><<<g, FB, 0, st>>>(bnd, f.X, f.G, f.np, f.cap, seg, nseg, part);
        break;
    case WINCKELMANS:
        ffv_partial_kernel<WINCKELMANS><<<g, FB, 0, st>>>(bnd, f.X, f.G, f.np, f.cap, seg, nseg, part);
        break;
    default:
        return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    ffv_reduce_kernel<<<(unsigned)((6 * nb + FB - 1) / FB), FB, 0, st>>>(part, nb, nseg, out);
    return hipGetLastError();
}

}  // namespace vpmx
