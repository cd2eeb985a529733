// a3d_probe.hip -- memory-pattern probe of the VTI stress sweep (measurement support).
//
// Replays the stress sweep's memory accesses with no stencil arithmetic, so the time of
// the real sweep can be set against the time its access pattern alone takes on the same
// device. The probe walks the padded grid exactly as the sweep does: workgroups of 64 x 4
// lanes own an x-y tile, stream z, and are dealt to the XCDs in contiguous tile runs; each
// plane reads U and V (default cache policy, the sweep's LDS-staged fields), Q at plane
// z+7 and w at plane z+3 (the sweep's z-window leads), P, and the 16-byte packed model, and
// writes P, Q and w in place; cells in the 20-cell PML of a direction also read and write
// that direction's pair of split fields (px/qx, py/qy, pz/qz). A plane barrier and one LDS
// exchange stand in for the sweep's tile staging. Patterns:
//   0  the same six read and three write streams over every lane of every row, separate
//      output arrays, no plane leads (a plain stream-mix ceiling);
//   1  the sweep's interior pattern (updated x range, in place, leads);
//   2  1 + the z-PML split pairs;  3  2 + the y-PML pairs;  4  3 + the x-PML pairs (the
//      whole pattern of the VTI stress sweep that also updates w, directional split).
#include "a3d_device.h"
#include "a3d_internal.h"
#include "../../include/acoustic3d.h"

namespace a3d {
namespace {

constexpr int PTY = 4;    // tile rows (the sweep's tuned shape at 400x400x300)
constexpr int PNP = 20;   // PML width

struct ProbeArgs {
    const float *u, *v, *q, *w, *p;
    const f32x4* m;
    float *po, *qo, *wo;          // outputs (pattern 0: separate arrays)
    float *sx0, *sx1, *sy0, *sy1, *sz0, *sz1;
    int ld, pl, nnx, nny, nnz, ntx, nty, ntiles, grid;
};

template <int PAT>
__global__ __launch_bounds__(TX* PTY) void k_probe_pattern(ProbeArgs a) {
    const int per = a.grid >> 3;
    const int t = (int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3);
    if (t >= a.ntiles) return;
    const int tix = t % a.ntx, tiy = t / a.ntx;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x = tix * TX + tx, y = MM + tiy * PTY + ty;
    const bool yin = y < a.nny - MM;
    const bool valid = PAT == 0 ? (yin && x < a.nnx) : (yin && x >= MM && x < a.nnx - MM);
    const int col = min(max(x, PAT == 0 ? 0 : MM), a.nnx - (PAT == 0 ? 1 : MM + 1)) + a.ld * min(y, a.nny - MM - 1);
    const bool dX = PAT >= 4 && (x < MM + PNP || x >= a.nnx - MM - PNP);
    const bool dY = PAT >= 3 && (y < MM + PNP || y >= a.nny - MM - PNP);
    __shared__ float s[PTY][TX];
    for (int z = MM; z < a.nnz - MM; ++z) {
        const int zo = z * a.pl;
        const int zq = PAT == 0 ? zo : min(z + 7, a.nnz - 1) * a.pl;
        const int zw = PAT == 0 ? zo : (z + 3) * a.pl;
        const bool dZ = PAT >= 2 && (z < MM + PNP || z >= a.nnz - MM - PNP);
        const float u = bld(a.u, zo, col), v = bld(a.v, zo, col);
        const float q = bld<NT>(a.q, zq, col), w = bld<NT>(a.w, zw, col), p = bld<NT>(a.p, zo, col);
        const f32x4 m = bld4<NT>(a.m, zo, col);
        float x0 = 0.f, x1 = 0.f, y0 = 0.f, y1 = 0.f, z0 = 0.f, z1 = 0.f;
        if (valid && dX) {
            x0 = bld<NT>(a.sx0, zo, col);
            x1 = bld<NT>(a.sx1, zo, col);
        }
        if (valid && dY) {
            y0 = bld<NT>(a.sy0, zo, col);
            y1 = bld<NT>(a.sy1, zo, col);
        }
        if (valid && dZ) {
            z0 = bld<NT>(a.sz0, zo, col);
            z1 = bld<NT>(a.sz1, zo, col);
        }
        __syncthreads();
        s[ty][tx] = u;
        __syncthreads();
        const float r = s[ty][(tx + 1) & (TX - 1)] + v + q + w + p + m.x + m.y + m.z + m.w;
        if (valid) {
            if (dX) {
                bst<NT>(a.sx0, zo, col, r + x0);
                bst<NT>(a.sx1, zo, col, r + x1);
            }
            if (dY) {
                bst<NT>(a.sy0, zo, col, r + y0);
                bst<NT>(a.sy1, zo, col, r + y1);
            }
            if (dZ) {
                bst<NT>(a.sz0, zo, col, r + z0);
                bst<NT>(a.sz1, zo, col, r + z1);
            }
            bst<NT>(a.po, zo, col, r);
            bst<NT>(a.qo, zo, col, r + 1.f);
            bst<NT>(a.wo, zw, col, r + 2.f);
        }
    }
}

hipError_t launch_probe_pattern(int pat, const ProbeArgs& a, hipStream_t st) {
    const dim3 grid(a.grid), block(TX, PTY);
    switch (pat) {
        case 0: hipLaunchKernelGGL(k_probe_pattern<0>, grid, block, 0, st, a); break;
        case 1: hipLaunchKernelGGL(k_probe_pattern<1>, grid, block, 0, st, a); break;
        case 2: hipLaunchKernelGGL(k_probe_pattern<2>, grid, block, 0, st, a); break;
        case 3: hipLaunchKernelGGL(k_probe_pattern<3>, grid, block, 0, st, a); break;
        default: hipLaunchKernelGGL(k_probe_pattern<4>, grid, block, 0, st, a); break;
    }
    return hipGetLastError();
}

__global__ void k_probe_sincos(const float* __restrict__ t, float* __restrict__ s, float* __restrict__ c, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rot_sincos(t[i], s[i], c[i]);
}

}  // namespace
}  // namespace a3d

using namespace a3d;

extern "C" int a3d_probe_rot_sincos(int device, const float* t, float* s, float* c, int n) {
    if (!t || !s || !c || n < 1) return report_error(A3D_EINVAL, "a3d_probe_rot_sincos", "bad argument");
    if (hipSetDevice(device) != hipSuccess) return report_error(A3D_EHIP, "a3d_probe_rot_sincos", "hipSetDevice");
    float* d = nullptr;
    if (hipMalloc(&d, 3 * (size_t)n * sizeof(float)) != hipSuccess)
        return report_error(A3D_ENOMEM, "a3d_probe_rot_sincos", "allocation failed");
    hipError_t e = hipMemcpy(d, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_probe_sincos, dim3((n + 255) / 256), dim3(256), 0, 0, d, d + n, d + 2 * (size_t)n, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(s, d + n, (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(c, d + 2 * (size_t)n, (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return report_error(A3D_EHIP, "a3d_probe_rot_sincos", hipGetErrorString(e));
    return A3D_OK;
}

extern "C" int a3d_probe_stress_pattern(int device, int nnx, int nny, int nnz, int reps, double* ms,
                                        double* bytes) {
    if (!ms || !bytes || reps < 1) return report_error(A3D_EINVAL, "a3d_probe_stress_pattern", "bad argument");
    if (nnx < 2 * (MM + PNP) + 1 || nny < 2 * (MM + PNP) + 1 || nnz < 2 * (MM + PNP) + 1)
        return report_error(A3D_EINVAL, "a3d_probe_stress_pattern", "grid smaller than the two PMLs");
    ProbeArgs a{};
    a.nnx = nnx;
    a.nny = nny;
    a.nnz = nnz;
    a.ld = (nnx + TX - 1) / TX * TX;
    if ((double)a.ld * nny * 16.0 >= 2147483647.0 || (double)a.ld * nny * nnz >= 2147483647.0)
        return report_error(A3D_EINVAL, "a3d_probe_stress_pattern", "grid exceeds the 31-bit offsets");
    a.pl = a.ld * nny;
    a.ntx = a.ld / TX;
    a.nty = (nny - 2 * MM + PTY - 1) / PTY;
    a.ntiles = a.ntx * a.nty;
    a.grid = (a.ntiles + 7) / 8 * 8;
    const size_t n = (size_t)a.pl * nnz;
    if (hipSetDevice(device) != hipSuccess) return report_error(A3D_EHIP, "a3d_probe_stress_pattern", "hipSetDevice");
    // 14 four-byte fields and the 16-byte model, zeroed (the probe stores finite sums)
    float* f[14] = {};
    f32x4* m = nullptr;
    int rc = A3D_OK;
    for (int i = 0; i < 14 && !rc; ++i)
        if (hipMalloc(&f[i], n * sizeof(float)) != hipSuccess || hipMemset(f[i], 0, n * sizeof(float)) != hipSuccess)
            rc = report_error(A3D_ENOMEM, "a3d_probe_stress_pattern", "field allocation failed");
    if (!rc && (hipMalloc(&m, n * sizeof(f32x4)) != hipSuccess || hipMemset(m, 0, n * sizeof(f32x4)) != hipSuccess))
        rc = report_error(A3D_ENOMEM, "a3d_probe_stress_pattern", "model allocation failed");
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc && (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = report_error(A3D_EHIP, "a3d_probe_stress_pattern", "stream/event creation failed");
    if (!rc) {
        a.u = f[0];
        a.v = f[1];
        a.q = f[2];
        a.w = f[3];
        a.p = f[4];
        a.m = m;
        a.sx0 = f[5];
        a.sx1 = f[6];
        a.sy0 = f[7];
        a.sy1 = f[8];
        a.sz0 = f[9];
        a.sz1 = f[10];
    }
    // algorithmic bytes per launch: 48 per updated point, 16 more per PML direction of a cell
    const double nxu = nnx - 2 * MM, nyu = nny - 2 * MM, nzu = nnz - 2 * MM;
    const double pts = nxu * nyu * nzu;
    const double extra[5] = {0.0, 0.0, 2.0 * PNP * nxu * nyu, 2.0 * PNP * nxu * nzu, 2.0 * PNP * nyu * nzu};
    double add = 0.0;
    for (int pat = 0; pat < 5 && !rc; ++pat) {
        add += extra[pat];
        ProbeArgs b = a;
        if (pat == 0) {
            b.po = f[11];
            b.qo = f[12];
            b.wo = f[13];
        } else {
            b.po = f[4];
            b.qo = f[2];
            b.wo = f[3];
        }
        auto timed = [&]() -> int {
            if (launch_probe_pattern(pat, b, st) != hipSuccess) return A3D_EHIP;   // warm-up
            if (hipEventRecord(e0, st) != hipSuccess) return A3D_EHIP;
            for (int r = 0; r < reps; ++r)
                if (launch_probe_pattern(pat, b, st) != hipSuccess) return A3D_EHIP;
            if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return A3D_EHIP;
            float t = 0.f;
            if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) return A3D_EHIP;
            ms[pat] = t / reps;
            return A3D_OK;
        };
        if (timed()) rc = report_error(A3D_EHIP, "a3d_probe_stress_pattern", "probe launch failed");
        const double pts0 = pat == 0 ? (double)nnx * nyu * nzu : pts;
        bytes[pat] = 48.0 * pts0 + 16.0 * add;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    for (float* p : f)
        if (p) (void)hipFree(p);
    if (m) (void)hipFree(m);
    return rc;
}
