/*
 * rf_oracle.c - CPU restatement of the reference's RF kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the gfx950
 * kernels of libgeraf_rf.so: only tests, the smoke check and the CPU-baseline
 * leg of the benchmark may load it.  No product code path calls it.
 *
 * Every function restates one reference kernel, in FP64 unless its name ends
 * in _ref32 (FP32 arithmetic in the reference's own operation order), with
 * the loops in the most direct order and OpenMP over independent outputs:
 *
 *   orc_mf_forward_f64      MatchedFilter_cuda_kernel.cu:65-115
 *   orc_mf_forward_ref32    MatchedFilter_cuda_kernel.cu:65-115 (FP32, per-term sinf/cosf)
 *   orc_mf_backward_f64     MatchedFilter_cuda_kernel.cu:171-220
 *   orc_st_forward_f64      SignalTracer_cuda_kernel.cu:351-450 (+ helpers :24-197)
 *   orc_st_backward_f64     SignalTracer_cuda_kernel.cu:452-587 (+ Jacobians :91-170)
 *   orc_bilinear_forward_f64  SignalTracer_cuda_kernel.cu:199-279
 *   orc_bilinear_backward_f64 SignalTracer_cuda_kernel.cu:281-349 (serial scatter)
 *
 * Scalar parameters use the reference's units: f0 [1e8 Hz], slope (As_sci)
 * [1e8 Hz/s], fs [Hz], c [1e8 m/s], path bias [m].
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define TWO_PI 6.283185307179586476925286766559

typedef struct {
  double f0, slope, fs, c, bias;
} orc_params;

/* ------------------------------------------------------------------ matched filter */

static double orc_dist(const double* a, double x, double y, double z) {
  const double dx = a[0] - x, dy = a[1] - y, dz = a[2] - z;
  return sqrt(dx * dx + dy * dy + dz * dz);
}

/* out[v] = (1/A) sum_a sum_k s[a,k] exp(-j phi), phi = 2pi (f0 + slope k/fs) tau,
 * tau = (|tx-v| + |rx-v| + bias)/c.  voxel is [3,N] (SoA). */
void orc_mf_forward_f64(const double* re, const double* im, const double* tx, const double* rx,
                        const double* voxel, double* out_re, double* out_im, int64_t A,
                        int64_t K, int64_t N, const orc_params* p) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t v = 0; v < N; ++v) {
    const double x = voxel[v], y = voxel[N + v], z = voxel[2 * N + v];
    double ar = 0.0, ai = 0.0;
    for (int64_t a = 0; a < A; ++a) {
      const double d = orc_dist(tx + 3 * a, x, y, z) + orc_dist(rx + 3 * a, x, y, z) + p->bias;
      const double tau = d / p->c;
      for (int64_t k = 0; k < K; ++k) {
        const double phase = TWO_PI * (p->f0 + p->slope * ((double)k / p->fs)) * tau;
        const double c = cos(phase), ns = -sin(phase);
        const double sr = re[a * K + k], si = im[a * K + k];
        ar += c * sr - ns * si;
        ai += c * si + ns * sr;
      }
    }
    out_re[v] = ar / (double)A;
    out_im[v] = ai / (double)A;
  }
}

/* The same operator in FP32 with the reference kernel's operation order. */
void orc_mf_forward_ref32(const float* re, const float* im, const float* tx, const float* rx,
                          const float* voxel, float* out_re, float* out_im, int64_t A,
                          int64_t K, int64_t N, const orc_params* p) {
  const float f0 = (float)p->f0, slope = (float)p->slope, fs = (float)p->fs, c = (float)p->c,
              bias = (float)p->bias;
  const float kPi = 3.14159265358979323846f;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t v = 0; v < N; ++v) {
    const float x = voxel[v], y = voxel[N + v], z = voxel[2 * N + v];
    float ar = 0.0f, ai = 0.0f;
    for (int64_t a = 0; a < A; ++a) {
      const float* t = tx + 3 * a;
      const float* r = rx + 3 * a;
      const float dt = sqrtf((t[0] - x) * (t[0] - x) + (t[1] - y) * (t[1] - y) +
                             (t[2] - z) * (t[2] - z));
      const float dr = sqrtf((r[0] - x) * (r[0] - x) + (r[1] - y) * (r[1] - y) +
                             (r[2] - z) * (r[2] - z));
      const float tau = (dt + dr + bias) / c;
      const float base = 2.0f * kPi * f0 * tau;
      for (int64_t k = 0; k < K; ++k) {
        const float st = (float)k / fs;
        const float phase = base + 2.0f * kPi * slope * st * tau;
        const float cs = cosf(phase), ns = -sinf(phase);
        const float sr = re[a * K + k], si = im[a * K + k];
        ar += cs * sr - ns * si;
        ai += cs * si + ns * sr;
      }
    }
    out_re[v] = ar / (float)A;
    out_im[v] = ai / (float)A;
  }
}

/* g[a,k] = (1/norm) sum_v gout[v] exp(+j phi_{a,k}(v)) (real/imag split as the
 * reference writes it: g_re = sum cos*gr - sin*gi, g_im = sum sin*gr + cos*gi). */
void orc_mf_backward_f64(const double* gr, const double* gi, const double* tx, const double* rx,
                         const double* voxel, double* g_re, double* g_im, int64_t A, int64_t K,
                         int64_t N, int64_t norm, const orc_params* p) {
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int64_t a = 0; a < A; ++a) {
    for (int64_t k = 0; k < K; ++k) {
      double ar = 0.0, ai = 0.0;
      const double st = (double)k / p->fs;
      for (int64_t v = 0; v < N; ++v) {
        const double x = voxel[v], y = voxel[N + v], z = voxel[2 * N + v];
        const double d = orc_dist(tx + 3 * a, x, y, z) + orc_dist(rx + 3 * a, x, y, z) + p->bias;
        const double tau = d / p->c;
        const double phase = TWO_PI * p->f0 * tau + TWO_PI * p->slope * st * tau;
        const double c = cos(phase), ns = -sin(phase);
        ar += c * gr[v] + ns * gi[v];
        ai += (-ns) * gr[v] + c * gi[v];
      }
      g_re[a * K + k] = ar / (double)norm;
      g_im[a * K + k] = ai / (double)norm;
    }
  }
}

/* ------------------------------------------------------------------ signal tracer */

#define ST_EPS 1.0e-7

static double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static void normalize3(const double* in, double* out) {
  const double inv = 1.0 / (sqrt(dot3(in, in)) + ST_EPS);
  out[0] = in[0] * inv;
  out[1] = in[1] * inv;
  out[2] = in[2] * inv;
}

typedef struct {
  double spec, idn;
  double inc[3], refl_raw[3], refl[3], ret[3];
} st_geom;

/* ComputeSpecularGeometry (SignalTracer_cuda_kernel.cu:172-197). */
static void st_geometry(const double* t, const double* r, const double* n, const double* pt,
                        st_geom* g) {
  double inc[3], ret[3];
  for (int i = 0; i < 3; ++i) {
    inc[i] = pt[i] - t[i];
    ret[i] = r[i] - pt[i];
  }
  normalize3(inc, g->inc);
  g->idn = dot3(g->inc, n);
  for (int i = 0; i < 3; ++i) g->refl_raw[i] = g->inc[i] - 2.0 * g->idn * n[i];
  normalize3(g->refl_raw, g->refl);
  normalize3(ret, g->ret);
  g->spec = dot3(g->ret, g->refl);
}

static double st_path(const double* t, const double* r, const double* pt) {
  return orc_dist(t, pt[0], pt[1], pt[2]) + orc_dist(r, pt[0], pt[1], pt[2]);
}

static int st_culled(const st_geom* g) { return g->spec < 1.0e-6 || g->idn > 0.0; }

/* out[a,k] = (1/P) sum_{p inbounds, not culled} sigma[a,p] spec / d^2 exp(+j phi). */
void orc_st_forward_f64(const double* normals, const double* sigmas, const double* points,
                        const double* tx, const double* rx, const uint8_t* inb, double* out_re,
                        double* out_im, int64_t A, int64_t P, int64_t K, const orc_params* p) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t a = 0; a < A; ++a) {
    double* ore = out_re + a * K;
    double* oim = out_im + a * K;
    for (int64_t k = 0; k < K; ++k) ore[k] = oim[k] = 0.0;
    for (int64_t q = 0; q < P; ++q) {
      if (!inb[q]) continue;
      st_geom g;
      st_geometry(tx + 3 * a, rx + 3 * a, normals + 3 * q, points + 3 * q, &g);
      if (st_culled(&g)) continue;
      const double d = st_path(tx + 3 * a, rx + 3 * a, points + 3 * q);
      const double amp = sigmas[a * P + q] * g.spec / (d * d);
      const double tau = (d + p->bias) / p->c;
      for (int64_t k = 0; k < K; ++k) {
        const double phase = (TWO_PI * p->f0 + TWO_PI * p->slope * ((double)k / p->fs)) * tau;
        ore[k] += amp * cos(phase);
        oim[k] += amp * sin(phase);
      }
    }
    for (int64_t k = 0; k < K; ++k) {
      ore[k] /= (double)P;
      oim[k] /= (double)P;
    }
  }
}

/* Gradients into normals [P,3] and sigmas [A,P].  Points that are out of
 * bounds get zero gradients (the reference leaves them unwritten). */
void orc_st_backward_f64(const double* g_re, const double* g_im, const double* normals,
                         const double* sigmas, const double* points, const double* tx,
                         const double* rx, const uint8_t* inb, double* grad_normals,
                         double* grad_sigmas, int64_t A, int64_t P, int64_t K,
                         const orc_params* p) {
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t q = 0; q < P; ++q) {
    double gn[3] = {0.0, 0.0, 0.0};
    const double* n = normals + 3 * q;
    for (int64_t a = 0; a < A; ++a) grad_sigmas[a * P + q] = 0.0;
    if (inb[q]) {
      for (int64_t a = 0; a < A; ++a) {
        st_geom g;
        st_geometry(tx + 3 * a, rx + 3 * a, n, points + 3 * q, &g);
        if (st_culled(&g)) continue;
        /* d refl_raw / d n  (ComputeReflectedDirectionNormalJacobian) */
        double J[9], NJ[9], gs[3];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c)
            J[r * 3 + c] = -2.0 * n[r] * g.inc[c] - (r == c ? 2.0 * g.idn : 0.0);
        /* (I - refl refl^T) J / (|refl_raw| + eps)  (ComputeNormalizedVectorJacobian) */
        const double inv = 1.0 / (sqrt(dot3(g.refl_raw, g.refl_raw)) + ST_EPS);
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int m = 0; m < 3; ++m)
              s += ((r == m ? 1.0 : 0.0) - g.refl[r] * g.refl[m]) * J[m * 3 + c];
            NJ[r * 3 + c] = s * inv;
          }
        /* d spec / d n = NJ^T ret  (ProjectDirectionJacobianOntoReturnVector) */
        for (int c = 0; c < 3; ++c)
          gs[c] = NJ[0 * 3 + c] * g.ret[0] + NJ[1 * 3 + c] * g.ret[1] + NJ[2 * 3 + c] * g.ret[2];
        const double d = st_path(tx + 3 * a, rx + 3 * a, points + 3 * q);
        const double att = 1.0 / (d * d);
        const double sv = sigmas[a * P + q];
        const double tau = (d + p->bias) / p->c;
        double acc = 0.0;
        for (int64_t k = 0; k < K; ++k) {
          const double phase = (TWO_PI * p->f0 + TWO_PI * p->slope * ((double)k / p->fs)) * tau;
          acc += g_re[a * K + k] * cos(phase) + g_im[a * K + k] * sin(phase);
        }
        for (int c = 0; c < 3; ++c) gn[c] += sv * att * gs[c] * acc;
        grad_sigmas[a * P + q] = att * g.spec * acc / (double)P;
      }
    }
    for (int c = 0; c < 3; ++c) grad_normals[3 * q + c] = gn[c] / (double)P;
  }
}

/* ------------------------------------------------------------------ bilinear resample */

static int64_t bl_lower(const double* z, int64_t S, double t) {
  int64_t lo = 0, hi = S - 1;
  while (lo < hi - 1) {
    const int64_t mid = (lo + hi) / 2;
    if (z[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

/* normals [R,S,3], sigmas [A,R,S], z [R,S], tz [R,T] -> out_n [R,T,3],
 * out_s [A,R,T], inb [R,T].  Requires S >= 2 (the reference reads z[1]). */
void orc_bilinear_forward_f64(const double* normals, const double* sigmas, const double* z,
                              const double* tz, double* out_n, double* out_s, uint8_t* inb,
                              int64_t A, int64_t R, int64_t S, int64_t T) {
#pragma omp parallel for collapse(2)
  for (int64_t a = 0; a < A; ++a) {
    for (int64_t r = 0; r < R; ++r) {
      const double* zr = z + r * S;
      for (int64_t t = 0; t < T; ++t) {
        const double tv = tz[r * T + t];
        const int64_t lo = bl_lower(zr, S, tv);
        const double z0 = zr[lo], z1 = zr[lo + 1];
        const int oob = (tv < z0) || (tv > z1);
        if (a == 0) inb[r * T + t] = (uint8_t)!oob;
        double* os = out_s + (a * R + r) * T + t;
        if (oob) {
          *os = 0.0;
          if (a == 0)
            for (int c = 0; c < 3; ++c) out_n[(r * T + t) * 3 + c] = 0.0;
          continue;
        }
        const double den = (z1 - z0) > 1.0e-6 ? (z1 - z0) : 1.0e-6;
        const double wu = (tv - z0) / den, wl = 1.0 - wu;
        const double* sr = sigmas + (a * R + r) * S;
        *os = wl * sr[lo] + wu * sr[lo + 1];
        if (a == 0)
          for (int c = 0; c < 3; ++c)
            out_n[(r * T + t) * 3 + c] =
                wl * normals[(r * S + lo) * 3 + c] + wu * normals[(r * S + lo + 1) * 3 + c];
      }
    }
  }
}

/* Transpose of the forward: scatter into grad_n [R,S,3] and grad_s [A,R,S]
 * (both zeroed here). */
void orc_bilinear_backward_f64(const double* g_n, const double* g_s, const double* z,
                               const double* tz, double* grad_n, double* grad_s, int64_t A,
                               int64_t R, int64_t S, int64_t T) {
  memset(grad_n, 0, sizeof(double) * (size_t)(R * S * 3));
  memset(grad_s, 0, sizeof(double) * (size_t)(A * R * S));
#pragma omp parallel for
  for (int64_t r = 0; r < R; ++r) {
    const double* zr = z + r * S;
    for (int64_t a = 0; a < A; ++a) {
      for (int64_t t = 0; t < T; ++t) {
        const double tv = tz[r * T + t];
        const int64_t lo = bl_lower(zr, S, tv);
        const double z0 = zr[lo], z1 = zr[lo + 1];
        if (tv < z0 || tv > z1) continue;
        const double den = (z1 - z0) > 1.0e-6 ? (z1 - z0) : 1.0e-6;
        const double wu = (tv - z0) / den, wl = 1.0 - wu;
        const double gs = g_s[(a * R + r) * T + t];
        grad_s[(a * R + r) * S + lo] += wl * gs;
        grad_s[(a * R + r) * S + lo + 1] += wu * gs;
        if (a == 0)
          for (int c = 0; c < 3; ++c) {
            const double gn = g_n[(r * T + t) * 3 + c];
            grad_n[(r * S + lo) * 3 + c] += wl * gn;
            grad_n[(r * S + lo + 1) * 3 + c] += wu * gn;
          }
      }
    }
  }
}

/* ------------------------------------------------------------------ range-bin matched filter */

/* MatchedFilter_cuda_kernel.cu:117-169 / :222-279.  d includes the bias; pairs
 * with d < 0 or d > max_d (the reference's FP32 fft_spacing*num_range_bins) are
 * skipped; bin = floor(d/spacing/2). */
static int64_t fft_bin(double d, double spacing, double max_d, int64_t R) {
  if (d < 0.0 || d > max_d) return -1;
  const double b = floor(d / spacing / 2.0);
  return (b < 0.0 || b >= (double)R) ? -1 : (int64_t)b;
}

void orc_mf_fft_forward_f64(const double* re, const double* im, const double* tx,
                            const double* rx, const double* voxel, double* out_re,
                            double* out_im, int64_t A, int64_t R, int64_t N, double spacing,
                            double max_d, const orc_params* p) {
#pragma omp parallel for
  for (int64_t v = 0; v < N; ++v) {
    const double x = voxel[v], y = voxel[N + v], z = voxel[2 * N + v];
    double ar = 0.0, ai = 0.0;
    for (int64_t a = 0; a < A; ++a) {
      const double d = orc_dist(tx + 3 * a, x, y, z) + orc_dist(rx + 3 * a, x, y, z) + p->bias;
      const int64_t b = fft_bin(d, spacing, max_d, R);
      if (b < 0) continue;
      const double phase = TWO_PI * p->f0 * (d / p->c);
      const double c = cos(phase), ns = -sin(phase);
      ar += c * re[a * R + b] - ns * im[a * R + b];
      ai += c * im[a * R + b] + ns * re[a * R + b];
    }
    out_re[v] = ar / (double)A;
    out_im[v] = ai / (double)A;
  }
}

void orc_mf_fft_backward_f64(const double* gr, const double* gi, const double* tx,
                             const double* rx, const double* voxel, double* g_re, double* g_im,
                             int64_t A, int64_t R, int64_t N, int64_t norm, double spacing,
                             double max_d, const orc_params* p) {
#pragma omp parallel for collapse(2)
  for (int64_t a = 0; a < A; ++a) {
    for (int64_t b = 0; b < R; ++b) {
      double ar = 0.0, ai = 0.0;
      for (int64_t v = 0; v < N; ++v) {
        const double x = voxel[v], y = voxel[N + v], z = voxel[2 * N + v];
        const double d =
            orc_dist(tx + 3 * a, x, y, z) + orc_dist(rx + 3 * a, x, y, z) + p->bias;
        if (fft_bin(d, spacing, max_d, R) != b) continue;
        const double phase = TWO_PI * p->f0 * (d / p->c);
        const double c = cos(phase), ns = -sin(phase);
        ar += c * gr[v] + ns * gi[v];
        ai += (-ns) * gr[v] + c * gi[v];
      }
      g_re[a * R + b] = ar / (double)norm;
      g_im[a * R + b] = ai / (double)norm;
    }
  }
}
