// H7: orientation-summed 2D intensity for an explicit q-vector list on gfx950.
//
// I[k] = sum_r | sum_el f_el(k) sum_{j in el} exp(i (R_r^T q_k) . x_j) |^2
// (CalcFunctions.cpp:1026-1062: the q vector is rotated by R^T instead of the
// atoms, 1020-1022 and 1033).
//
// One thread per (q point, rotation); the atoms stream through LDS in chunks
// of 1024 (FP64 SoA), read by every lane at the same address (broadcast).
// The phase q'.x is formed in FP64 (in turns) and reduced to [0, 1), so the
// FP32 hardware sine and cosine (v_sin_f32 / v_cos_f32 take turns) see an
// argument accurate to 2^-24 turn even where q'.x reaches 10^2 turns; the
// element amplitudes are accumulated in FP64 and |A|^2 is added into the
// FP64 result with atomics (one per rotation and q point).
#include <algorithm>
#include <cmath>

#include "xs_internal.h"

namespace xs {
namespace amp2d {

constexpr int kBlock = 256;
constexpr int kChunk = 1024;
constexpr double kInvTwoPi = 0.15915494309189533576888376337251;

struct Args {
  const double *x, *y, *z;
  const double *qvec;  // nqp x 3
  const double *rot;   // nrot x 9, row-major R
  const double *ff;    // n_el x nqp
  double *I;
  unsigned nqp, n_el;
  unsigned el_off[17];
};

__global__ __launch_bounds__(kBlock) void amp2d_kernel(Args A) {
  __shared__ double sx[kChunk], sy[kChunk], sz[kChunk];
  const unsigned k = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = k < A.nqp;
  const double *R = A.rot + 9 * (size_t)blockIdx.y;
  double qx = 0, qy = 0, qz = 0;
  if (valid) {
    const double a = A.qvec[3 * (size_t)k], b = A.qvec[3 * (size_t)k + 1],
                 c = A.qvec[3 * (size_t)k + 2];
    qx = (a * R[0] + b * R[3] + c * R[6]) * kInvTwoPi;
    qy = (a * R[1] + b * R[4] + c * R[7]) * kInvTwoPi;
    qz = (a * R[2] + b * R[5] + c * R[8]) * kInvTwoPi;
  }
  double Ar = 0.0, Ai = 0.0;
  for (unsigned e = 0; e < A.n_el; ++e) {
    double ar = 0.0, ai = 0.0;
    const unsigned e0 = A.el_off[e], e1 = A.el_off[e + 1];
    for (unsigned c0 = e0; c0 < e1; c0 += kChunk) {
      const unsigned cn = min((unsigned)kChunk, e1 - c0);
      __syncthreads();
      for (unsigned j = threadIdx.x; j < cn; j += kBlock) {
        sx[j] = A.x[c0 + j];
        sy[j] = A.y[c0 + j];
        sz[j] = A.z[c0 + j];
      }
      __syncthreads();
#pragma unroll 4
      for (unsigned j = 0; j < cn; ++j) {
        const double ph = fma(qx, sx[j], fma(qy, sy[j], qz * sz[j]));
        const float f = (float)(ph - floor(ph));
        ar += (double)__builtin_amdgcn_cosf(f);
        ai += (double)__builtin_amdgcn_sinf(f);
      }
    }
    if (valid) {
      const double w = A.ff[(size_t)e * A.nqp + k];
      Ar = fma(w, ar, Ar);
      Ai = fma(w, ai, Ai);
    }
  }
  if (valid) atomicAdd(&A.I[k], Ar * Ar + Ai * Ai);
}

// Regular q grid: q(iu, iv) = q0 + iu du + iv dv, point k = iv nu + iu.
// One thread per run of kRun consecutive points of a row and one rotation.
// For an atom x the phase advances by delta = (R^T du).x per point, so the
// run's terms are z_m = exp(i ph0) e^{i m delta}: z_0 from one FP32 hardware
// sine/cosine of the FP64-reduced first phase (as the list kernel), z_1 =
// z_0 (cos delta + i sin delta), then z_{m+1} = 2 cos(delta) z_m - z_{m-1}
// (two FP64 FMAs a point).  cos and sin of delta are computed once per atom
// and work group in FP64 and kept in LDS.  The three-term recurrence scales
// a rounding error made at step k by |sin((m-k) delta) / sin(delta)| <= m - k,
// so over a run of 8 it adds < 1e-14 to the seed's error (2^-24 turn).
constexpr int kRun = 16;
constexpr int kGridChunk = 512;

struct GridArgs {
  const double *x, *y, *z;
  const double *rot;  // nrot x 9
  const double *ff;   // n_el x nqp
  double *I;
  double q0[3], du[3], dv[3];
  unsigned nu, nv, runs_per_row, n_el;
  unsigned el_off[17];
};

__global__ __launch_bounds__(kBlock) void amp2d_grid_kernel(GridArgs A) {
  __shared__ double sx[kGridChunk], sy[kGridChunk], sz[kGridChunk], sc[kGridChunk],
      ss[kGridChunk];
  const unsigned run = blockIdx.x * kBlock + threadIdx.x;
  const unsigned iv = run / A.runs_per_row;
  const unsigned iu0 = (run - iv * A.runs_per_row) * kRun;
  const bool valid = iv < A.nv;
  const double *R = A.rot + 9 * (size_t)blockIdx.y;
  // R^T applied to a vector, in turns
  auto turn = [&](double a, double b, double c, double *o) {
    o[0] = (a * R[0] + b * R[3] + c * R[6]) * kInvTwoPi;
    o[1] = (a * R[1] + b * R[4] + c * R[7]) * kInvTwoPi;
    o[2] = (a * R[2] + b * R[5] + c * R[8]) * kInvTwoPi;
  };
  double q[3] = {0, 0, 0}, d[3];
  if (valid) {
    const double u = (double)iu0, v = (double)iv;
    turn(A.q0[0] + u * A.du[0] + v * A.dv[0], A.q0[1] + u * A.du[1] + v * A.dv[1],
         A.q0[2] + u * A.du[2] + v * A.dv[2], q);
  }
  turn(A.du[0], A.du[1], A.du[2], d);
  const size_t k0 = (size_t)iv * A.nu + iu0;
  double Ar[kRun], Ai[kRun];
#pragma unroll
  for (int m = 0; m < kRun; ++m) Ar[m] = Ai[m] = 0.0;
  for (unsigned e = 0; e < A.n_el; ++e) {
    double ar[kRun], ai[kRun];
#pragma unroll
    for (int m = 0; m < kRun; ++m) ar[m] = ai[m] = 0.0;
    const unsigned e0 = A.el_off[e], e1 = A.el_off[e + 1];
    for (unsigned c0 = e0; c0 < e1; c0 += kGridChunk) {
      const unsigned cn = min((unsigned)kGridChunk, e1 - c0);
      __syncthreads();
      for (unsigned j = threadIdx.x; j < cn; j += kBlock) {
        const double xj = A.x[c0 + j], yj = A.y[c0 + j], zj = A.z[c0 + j];
        sx[j] = xj;
        sy[j] = yj;
        sz[j] = zj;
        const double st = fma(d[0], xj, fma(d[1], yj, d[2] * zj));
        double sn, cs;
        sincospi(2.0 * (st - floor(st)), &sn, &cs);
        sc[j] = cs;
        ss[j] = sn;
      }
      __syncthreads();
      for (unsigned j = 0; j < cn; ++j) {
        const double ph = fma(q[0], sx[j], fma(q[1], sy[j], q[2] * sz[j]));
        const float f = (float)(ph - floor(ph));
        const double c = sc[j], sn = ss[j], c2 = 2.0 * c;
        double rp = (double)__builtin_amdgcn_cosf(f), ip = (double)__builtin_amdgcn_sinf(f);
        double re = fma(rp, c, -ip * sn), im = fma(rp, sn, ip * c);
        ar[0] += rp;
        ai[0] += ip;
#pragma unroll
        for (int m = 1; m < kRun; ++m) {
          ar[m] += re;
          ai[m] += im;
          if (m + 1 < kRun) {
            const double rn = fma(c2, re, -rp), in = fma(c2, im, -ip);
            rp = re;
            ip = im;
            re = rn;
            im = in;
          }
        }
      }
    }
    if (valid) {
#pragma unroll
      for (int m = 0; m < kRun; ++m)
        if (iu0 + m < A.nu) {
          const double w = A.ff[(size_t)e * A.nu * A.nv + k0 + m];
          Ar[m] = fma(w, ar[m], Ar[m]);
          Ai[m] = fma(w, ai[m], Ai[m]);
        }
    }
  }
  if (valid) {
#pragma unroll
    for (int m = 0; m < kRun; ++m)
      if (iu0 + m < A.nu) atomicAdd(&A.I[k0 + m], Ar[m] * Ar[m] + Ai[m] * Ai[m]);
  }
}

__global__ void zero_f64(double *__restrict__ p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

}  // namespace amp2d
}  // namespace xs

using namespace xs;

extern "C" int xs_amp2d(xs_ctx *ctx, const double *xyz, size_t n_atoms, const unsigned *el_count,
                        unsigned n_el, const double *qvec, size_t nqp, const double *rot,
                        unsigned nrot, const double *ff, double *I, unsigned flags) {
  XS_TRY(activate(ctx));
  if (!el_count || !qvec || !rot || !ff || !I || (!xyz && n_atoms))
    return fail(XS_ERR_ARG, "xs_amp2d: null pointer");
  if (n_el == 0 || n_el > 16) return fail(XS_ERR_RANGE, "xs_amp2d: n_el must be in [1, 16]");
  size_t tot = 0;
  for (unsigned e = 0; e < n_el; ++e) tot += el_count[e];
  if (tot != n_atoms) return fail(XS_ERR_ARG, "xs_amp2d: sum(el_count) != n_atoms");
  if (n_atoms >= (size_t(1) << 32) || nqp >= (size_t(1) << 31))
    return fail(XS_ERR_RANGE, "xs_amp2d: size too large");
  if (nqp == 0) return XS_OK;
  if (nrot == 0 || nrot > 65535) return fail(XS_ERR_RANGE, "xs_amp2d: nrot must be in [1, 65535]");
  const bool device = flags & XS_DEVICE_PTRS;
  amp2d::Args A{};
  A.nqp = (unsigned)nqp;
  A.n_el = n_el;
  A.el_off[0] = 0;
  for (unsigned e = 0; e < n_el; ++e) A.el_off[e + 1] = A.el_off[e] + el_count[e];
  const double *xyzd = xyz, *qd = qvec, *rd = rot, *fd = ff;
  double *Id = I;
  if (!device) {
    XS_TRY(ensure_workspace(ctx, Carver::need({3 * n_atoms * 8, 3 * nqp * 8, (size_t)nrot * 72,
                                               (size_t)n_el * nqp * 8, nqp * 8})));
    Carver cv(ctx->ws);
    double *x2 = cv.take<double>(3 * n_atoms), *q2 = cv.take<double>(3 * nqp),
           *r2 = cv.take<double>((size_t)nrot * 9), *f2 = cv.take<double>((size_t)n_el * nqp);
    Id = cv.take<double>(nqp);
    XS_HIP_CHECK(hipMemcpyAsync(x2, xyz, 3 * n_atoms * 8, hipMemcpyHostToDevice, ctx->stream));
    XS_HIP_CHECK(hipMemcpyAsync(q2, qvec, 3 * nqp * 8, hipMemcpyHostToDevice, ctx->stream));
    XS_HIP_CHECK(hipMemcpyAsync(r2, rot, (size_t)nrot * 72, hipMemcpyHostToDevice, ctx->stream));
    XS_HIP_CHECK(hipMemcpyAsync(f2, ff, (size_t)n_el * nqp * 8, hipMemcpyHostToDevice, ctx->stream));
    xyzd = x2;
    qd = q2;
    rd = r2;
    fd = f2;
  }
  A.x = xyzd;
  A.y = xyzd + n_atoms;
  A.z = xyzd + 2 * n_atoms;
  A.qvec = qd;
  A.rot = rd;
  A.ff = fd;
  A.I = Id;
  const unsigned zb = (unsigned)std::min<size_t>((nqp + 255) / 256, 4096);
  hipLaunchKernelGGL(amp2d::zero_f64, dim3(zb), dim3(256), 0, ctx->stream, Id, nqp);
  hipEvent_t ev;
  XS_TRY(prof_begin(ctx, "amp2d", &ev));
  hipLaunchKernelGGL(amp2d::amp2d_kernel, dim3((unsigned)((nqp + amp2d::kBlock - 1) / amp2d::kBlock), nrot),
                     dim3(amp2d::kBlock), 0, ctx->stream, A);
  XS_HIP_CHECK(hipGetLastError());
  XS_TRY(prof_end(ctx, "amp2d", ev));
  if (!device) {
    XS_HIP_CHECK(hipMemcpyAsync(I, Id, nqp * 8, hipMemcpyDeviceToHost, ctx->stream));
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }
  return XS_OK;
}

extern "C" int xs_amp2d_grid(xs_ctx *ctx, const double *xyz, size_t n_atoms,
                             const unsigned *el_count, unsigned n_el, const double *q0,
                             const double *du, const double *dv, unsigned nu, unsigned nv,
                             const double *rot, unsigned nrot, const double *ff, double *I,
                             unsigned flags) {
  XS_TRY(activate(ctx));
  if (!el_count || !q0 || !du || !dv || !rot || !ff || !I || (!xyz && n_atoms))
    return fail(XS_ERR_ARG, "xs_amp2d_grid: null pointer");
  if (n_el == 0 || n_el > 16) return fail(XS_ERR_RANGE, "xs_amp2d_grid: n_el must be in [1, 16]");
  size_t tot = 0;
  for (unsigned e = 0; e < n_el; ++e) tot += el_count[e];
  if (tot != n_atoms) return fail(XS_ERR_ARG, "xs_amp2d_grid: sum(el_count) != n_atoms");
  const size_t nqp = (size_t)nu * nv;
  if (n_atoms >= (size_t(1) << 32) || nqp >= (size_t(1) << 31))
    return fail(XS_ERR_RANGE, "xs_amp2d_grid: size too large");
  if (nqp == 0) return XS_OK;
  if (nrot == 0 || nrot > 65535)
    return fail(XS_ERR_RANGE, "xs_amp2d_grid: nrot must be in [1, 65535]");
  const bool device = flags & XS_DEVICE_PTRS;
  amp2d::GridArgs A{};
  A.nu = nu;
  A.nv = nv;
  A.runs_per_row = (nu + amp2d::kRun - 1) / amp2d::kRun;
  A.n_el = n_el;
  for (int k = 0; k < 3; ++k) {
    A.q0[k] = q0[k];
    A.du[k] = du[k];
    A.dv[k] = dv[k];
  }
  A.el_off[0] = 0;
  for (unsigned e = 0; e < n_el; ++e) A.el_off[e + 1] = A.el_off[e] + el_count[e];
  const double *xyzd = xyz, *rd = rot, *fd = ff;
  double *Id = I;
  if (!device) {
    XS_TRY(ensure_workspace(ctx, Carver::need({3 * n_atoms * 8, (size_t)nrot * 72,
                                               (size_t)n_el * nqp * 8, nqp * 8})));
    Carver cv(ctx->ws);
    double *x2 = cv.take<double>(3 * n_atoms), *r2 = cv.take<double>((size_t)nrot * 9),
           *f2 = cv.take<double>((size_t)n_el * nqp);
    Id = cv.take<double>(nqp);
    XS_HIP_CHECK(hipMemcpyAsync(x2, xyz, 3 * n_atoms * 8, hipMemcpyHostToDevice, ctx->stream));
    XS_HIP_CHECK(hipMemcpyAsync(r2, rot, (size_t)nrot * 72, hipMemcpyHostToDevice, ctx->stream));
    XS_HIP_CHECK(hipMemcpyAsync(f2, ff, (size_t)n_el * nqp * 8, hipMemcpyHostToDevice, ctx->stream));
    xyzd = x2;
    rd = r2;
    fd = f2;
  }
  A.x = xyzd;
  A.y = xyzd + n_atoms;
  A.z = xyzd + 2 * n_atoms;
  A.rot = rd;
  A.ff = fd;
  A.I = Id;
  const unsigned zb = (unsigned)std::min<size_t>((nqp + 255) / 256, 4096);
  hipLaunchKernelGGL(amp2d::zero_f64, dim3(zb), dim3(256), 0, ctx->stream, Id, nqp);
  const size_t runs = (size_t)A.runs_per_row * nv;
  hipEvent_t ev;
  XS_TRY(prof_begin(ctx, "amp2d", &ev));
  hipLaunchKernelGGL(amp2d::amp2d_grid_kernel,
                     dim3((unsigned)((runs + amp2d::kBlock - 1) / amp2d::kBlock), nrot),
                     dim3(amp2d::kBlock), 0, ctx->stream, A);
  XS_HIP_CHECK(hipGetLastError());
  XS_TRY(prof_end(ctx, "amp2d", ev));
  if (!device) {
    XS_HIP_CHECK(hipMemcpyAsync(I, Id, nqp * 8, hipMemcpyDeviceToHost, ctx->stream));
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }
  return XS_OK;
}
