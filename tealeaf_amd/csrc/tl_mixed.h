// Internal interface of the mixed-precision kernels (FP32 sweeps, FP64
// dots and updates) used by the native driver.  FP32 arrays have row pitch
// P (a multiple of 4, >= x); FP64 arrays have row pitch x.  A 3D field is z
// planes of y such rows (z <= 1: 2D); the pointwise kernels take z.
// Scalars come back synchronously, like the reference's reductions.
#pragma once

namespace tl {
namespace mix {

// a[0..n) = 0
int zero32(float* a, long n);
// dst[row*P+col] = (float)src[row*x+col] over the whole chunk; pad columns 0.
int demote(const double* src, int x, int y, float* dst, int P);
// w = A p; *pw = sum w*p (FP64)
int calc_w32(int x, int y, int hd, int P, double* pw, const float* p, float* w, const float* kx,
             const float* ky);
// p_out = beta p_in + r, then w = A p_out; *pw = sum w*p_out (FP64)
int calc_pw32(int x, int y, int hd, int P, double beta, double* pw, const float* p_in,
              float* p_out, const float* r, float* w, const float* kx, const float* ky);
// d += alpha p; r -= alpha w; *rrn = sum r*r (FP64)
int calc_ur32(int x, int y, int z, int hd, int P, double alpha, double* rrn, float* d,
              const float* p, float* r, const float* w);
// p = beta p + r, or p = r when init
int calc_p32(int x, int y, int z, int hd, int P, double beta, bool init, float* p,
             const float* r);
// *out = sum a*a (FP64)
int norm32(int x, int y, int z, int hd, int P, const float* a, double* out);
// u += d (FP64); d = 0
int promote(int x, int y, int z, int hd, int P, double* u, float* d);
// calc_w32 (update false) or calc_pw32 (true) with *alpha_dev = rro / pw
// written on the device and no wait for the result; p_out's halo is not
// written (no kernel here reads an FP32 halo)
int calc_pw32_alpha(int x, int y, int hd, int P, bool update, double beta, double rro,
                    double* alpha_dev, const float* p_in, float* p_out, const float* r, float* w,
                    const float* kx, const float* ky);
// tlhip2d_ppcg_inner_iterations_f32, with cg_calc_ur's alpha read from the
// device (alpha_dev, when not null), and, when rrn is not null, also forming
// *rrn = sum r*r (FP64) over the r it writes in the same launch; the
// outputs' halos are not written
int ppcg_inner32_norm(int x, int y, int hd, int pitch, int count, const double* alphas,
                      const double* betas, double theta, int init_from_r, const int* neighbours,
                      const float* sd_in, float* sd_out, const float* r_in, float* r_out,
                      const float* u_in, float* u_out, const float* kx, const float* ky,
                      double alpha, const float* p, const float* w, const double* alpha_dev,
                      double* rrn);
// Chebyshev iterations on the correction equation A d = b (FP32): count
// turns of r = b - A d (d mirrored at the interior's edges), p = alpha_k p +
// beta_k r, d += p in one fused launch; when rrn is not null, *rrn = sum r*r
// (FP64) of the last turn's r.  Outputs' halos are not written.
int cheby_iterations32(int x, int y, int hd, int P, int count, const double* alphas,
                       const double* betas, const int* neighbours, const float* d_in,
                       float* d_out, const float* b, const float* p_in, float* p_out, float* r,
                       const float* kx, const float* ky, double* rrn);
// Chebyshev start on the correction equation (FP32): p = r / theta, d += p
int cheby_init32(int x, int y, int z, int hd, int P, double theta, float* p, const float* r,
                 float* d);
// r32 = (float)(u0 - A u) computed in FP64; *rr = sum (u0 - A u)^2 (FP64)
int residual64(int x, int y, int hd, int P, const double* u, const double* u0, const double* kx,
               const double* ky, float* r32, double* rr);

// 3D (7-point operator, neighbours outside the interior mirrored):
// w = A p, *pw = sum w*p (FP64)
int calc_w3_32(int x, int y, int z, int hd, int P, double* pw, const float* p, float* w,
               const float* kx, const float* ky, const float* kz);
// sd = r / theta
int ppcg_init3_32(int x, int y, int z, int hd, int P, double theta, const float* r, float* sd);
// one PPCG inner step: w = A sd_in, r -= w, d += sd_in, sd_out = alpha sd_in + beta r
// Two FP32 PPCG inner steps per pass (tl_kernels3d.hip): bit-identical to
// two ppcg_inner3_32 turns; r, d, sd from *_in to *_out; even x.
int ppcg_pair3_32(int x, int y, int z, int hd, int P, const double* alphas,
                  const double* betas, const float* sd_in, float* sd_out, const float* r_in,
                  float* r_out, const float* d_in, float* d_out, const float* kx,
                  const float* ky, const float* kz);
int ppcg_inner3_32(int x, int y, int z, int hd, int P, double alpha, double beta,
                   const float* sd_in, float* sd_out, float* r, float* d, const float* kx,
                   const float* ky, const float* kz);
// one Chebyshev turn on A d = b: r = b - A d_in, p = alpha p + beta r,
// d_out = d_in + p; *rr = sum r*r (FP64) when rr is not null
int cheby_iter3_32(int x, int y, int z, int hd, int P, double alpha, double beta,
                   const float* d_in, float* d_out, const float* b, float* p, float* r,
                   const float* kx, const float* ky, const float* kz, double* rr);
// r32 = (float)(u0 - A u) computed in FP64; *rr = sum (u0 - A u)^2 (FP64)
int residual3_64(int x, int y, int z, int hd, int P, const double* u, const double* u0,
                 const double* kx, const double* ky, const double* kz, float* r32, double* rr);

}  // namespace mix
}  // namespace tl
