/*
 * oak_oracle.c -- CPU restatement of OAK's local ensemble square-root analysis.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product path
 * (oak_amd/csrc).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only to compute expected values or the CPU
 * baseline time.  Nothing in the product links or calls it.
 *
 * Everything is FP64 ("PRECISION=double" build of OAK, Compilers/Linux-gfortran.mk:28-30)
 * and follows the operation order of the Fortran it cites, so that the results
 * agree with OAK to rounding.  Arrays are column-major (Fortran layout) with
 * explicit leading dimensions; indices are 0-based.
 *
 * Pinning (the tests that make this restatement trustworthy):
 *   oko_distance / oko_select   closed-form distances on hand-built geometries,
 *                               inclusive cutoff, three metrics evaluated in numpy
 *   oko_locfun                  known answers of test/test_covariance.F90:612-617
 *   oko_symeig (both backends)  numpy.linalg.eigh (LAPACK) + A·V = V·diag(E)
 *   oko_perp_space / rotate     Ω·w = v, ΩᵀΩ = I, P(u)ᵀu = 0
 *   oko_analysis_increment      textbook Kalman filter (test/test_rrsqrt.F90:27-77)
 *   oko_loc_analysis_*          single zone = global KF, one zone per element =
 *                               global KF, localised zones = dense KF with
 *                               R_loc⁻¹ = R⁻¹∘wwᵀ (test/test_rrsqrt.F90:142-236)
 *   oko_obsoper                 dense H·x brute force
 *   oko_ensemble_analysis       mean(Ea)=xa, cov(Ea)=Sa·Saᵀ, empty zones = identity
 *   oko_report                  hand-computed known answer; scaling of d; an
 *                               excluded observation = a removed one; sums over
 *                               observation ranges add up to the whole
 *   oko_ensemble_report         projection^2 ingrid = r^T cov(HE) r with numpy's
 *                               sample covariance, d = yo - mean(HE)
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OKO_OK 0
#define OKO_EARG 1
#define OKO_ENAN 2
#define OKO_ENOMEM 4

/* ------------------------------------------------------------------------- */
/* LAPACK dsyev / BLAS dgemm, resolved at run time (the reference calls them  */
/* through matoper_inc.F90:435,986-1005).  When no LAPACK is found the own    */
/* Jacobi solver below is used instead.                                       */
/* ------------------------------------------------------------------------- */
typedef void (*dsyev_fn)(const char*, const char*, const int*, double*, const int*,
                         double*, double*, const int*, int*, size_t, size_t);
typedef void (*dgemm_fn)(const char*, const char*, const int*, const int*, const int*,
                         const double*, const double*, const int*, const double*, const int*,
                         const double*, double*, const int*, size_t, size_t);
typedef void (*setnt_fn)(int);
static dsyev_fn p_dsyev = 0;
static dgemm_fn p_dgemm = 0;
static int eig_backend = 0; /* 0 = own Jacobi, 1 = LAPACK dsyev */
static int gemm_backend = 0; /* 0 = loops, 1 = BLAS dgemm */

/* path: shared library exporting dsyev_/dgemm_ (optionally with a prefix). */
int oko_load_lapack(const char* path, const char* prefix) {
  void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) return 1;
  char name[128];
  snprintf(name, sizeof name, "%sdsyev_", prefix);
  p_dsyev = (dsyev_fn)dlsym(h, name);
  snprintf(name, sizeof name, "%sdgemm_", prefix);
  p_dgemm = (dgemm_fn)dlsym(h, name);
  snprintf(name, sizeof name, "%sopenblas_set_num_threads", prefix);
  setnt_fn st = (setnt_fn)dlsym(h, name);
  if (!st) st = (setnt_fn)dlsym(h, "openblas_set_num_threads");
  if (st) st(1); /* one BLAS thread: zones are the parallel unit (rrsqrt.F90:357) */
  return (p_dsyev && p_dgemm) ? 0 : 2;
}

/* 0 = own Jacobi / loops, 1 = LAPACK dsyev + BLAS dgemm. */
int oko_set_backend(int b) {
  if (b == 1 && !(p_dsyev && p_dgemm)) return 1;
  eig_backend = b;
  gemm_backend = b;
  return 0;
}

int oko_get_backend(void) { return eig_backend; }

/* ------------------------------------------------------------------------- */
/* Distances and observation selection                                        */
/* ------------------------------------------------------------------------- */
static const double OKO_PI = 3.141592653589793238462643383279502884197;
static const double OKO_EARTH_RADIUS = 6378137.0;

/* assimilation.F90:3635-3672.  p0 = observation (lon/x, lat/y), p1 = zone.
 * metric 0 = Cartesian, 1 = spherical (great circle), 2 = spherical approx. */
double oko_distance(int metric, double p0x, double p0y, double p1x, double p1y) {
  if (metric == 0) {
    double dx = p1x - p0x, dy = p1y - p0y;
    double s = dx * dx;
    s = s + dy * dy;
    return sqrt(s);
  } else if (metric == 2) {
    double coeff = OKO_PI * OKO_EARTH_RADIUS / 180.;
    double a = coeff * cos((p0y + p1y) * (OKO_PI / 360.)) * (p1x - p0x);
    double b = coeff * (p1y - p0y);
    double s = a * a;
    s = s + b * b;
    return sqrt(s);
  } else {
    double d2r = OKO_PI / 180.;
    double a = p0y * d2r, b = p1y * d2r, C = (p1x - p0x) * d2r;
    double coeff = sin(b) * sin(a) + cos(b) * cos(a) * cos(C);
    coeff = fmax(fmin(coeff, 1.), -1.);
    return OKO_EARTH_RADIUS * acos(coeff);
  }
}

/* assimilation.F90:3683-3771 (selectObservations) for one zone whose first
 * element sits at zpos = (x, y, z, t) with correlation length L and cutoff
 * (maxLength).  loctype 1 = horizontal distance, 2 = |Δz|, 3 = |Δt|.
 * weight[l] = exp(-(d/L)^2) for all l when any observation is relevant
 * (otherwise weight holds the distances), relevant[l] = d <= cutoff.
 * Returns the number of relevant observations. */
int64_t oko_select(int metric, int loctype, const double* zpos, double L, double cutoff,
                   int64_t m, const double* ox, const double* oy, const double* oz,
                   const double* ot, double* weight, unsigned char* relevant) {
  int64_t cnt = 0;
  for (int64_t l = 0; l < m; ++l) {
    double d;
    if (loctype == 1)
      d = oko_distance(metric, ox[l], oy[l], zpos[0], zpos[1]);
    else if (loctype == 2)
      d = fabs(oz[l] - zpos[2]);
    else
      d = fabs(ot[l] - zpos[3]);
    weight[l] = d;
    relevant[l] = d <= cutoff;
    cnt += relevant[l];
  }
  if (cnt > 0)
    for (int64_t l = 0; l < m; ++l) {
      double q = weight[l] / L;
      weight[l] = exp(-(q * q));
    }
  return cnt;
}

/* covariance.F90:645-667 (Gaspari-Cohn locfun), used by test/test_rrsqrt.F90:254-271. */
double oko_locfun(double r) {
  if (r <= 1.)
    return (((-r / 4. + 1. / 2.) * r + 5. / 8.) * r - 5. / 3.) * (r * r) + 1.;
  if (r <= 2.)
    return ((((r / 12. - 1. / 2.) * r + 5. / 8.) * r + 5. / 3.) * r - 5.) * r + 4. - 2. / (3. * r);
  return 0.;
}

/* ------------------------------------------------------------------------- */
/* Symmetric eigensolver: matoper_inc.F90:939-1005 -> dsyev('V','L').          */
/* Eigenvalues ascending, eigenvectors in the columns of V (n x n, ld n).      */
/* The own backend is a cyclic two-sided Jacobi method (Jacobi 1846; Golub &   */
/* Van Loan, Matrix Computations, 4th ed., section 8.5, with Rutishauser's      */
/* stable rotation formulas), reading the lower triangle only.                */
/* ------------------------------------------------------------------------- */
static int jacobi_eig(int n, const double* A, int lda, double* E, double* V) {
  double* a = (double*)malloc(sizeof(double) * (size_t)n * n);
  if (!a) return OKO_ENOMEM;
  for (int j = 0; j < n; ++j)
    for (int i = j; i < n; ++i) {
      a[i + (size_t)j * n] = A[i + (size_t)j * lda];
      a[j + (size_t)i * n] = A[i + (size_t)j * lda];
    }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) V[i + (size_t)j * n] = (i == j);
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0, diag = 0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double x = a[i + (size_t)j * n];
        if (i == j) diag += x * x; else off += x * x;
      }
    if (off == 0 || off <= 1e-34 * diag) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = a[p + (size_t)q * n];
        double app = a[p + (size_t)p * n], aqq = a[q + (size_t)q * n];
        if (apq == 0) continue;
        if (sweep > 3 && fabs(apq) * 1e3 + fabs(app) == fabs(app) &&
            fabs(apq) * 1e3 + fabs(aqq) == fabs(aqq)) {
          a[p + (size_t)q * n] = a[q + (size_t)p * n] = 0;
          continue;
        }
        double theta = (aqq - app) / (2 * apq);
        double t = (theta >= 0 ? 1. : -1.) / (fabs(theta) + sqrt(theta * theta + 1.));
        double c = 1. / sqrt(t * t + 1.), s = t * c, tau = s / (1. + c);
        /* rows/cols p,q of a <- J^T a J */
        for (int k = 0; k < n; ++k) {
          double akp = a[k + (size_t)p * n], akq = a[k + (size_t)q * n];
          a[k + (size_t)p * n] = akp - s * (akq + tau * akp);
          a[k + (size_t)q * n] = akq + s * (akp - tau * akq);
        }
        for (int k = 0; k < n; ++k) {
          double apk = a[p + (size_t)k * n], aqk = a[q + (size_t)k * n];
          a[p + (size_t)k * n] = apk - s * (aqk + tau * apk);
          a[q + (size_t)k * n] = aqk + s * (apk - tau * aqk);
        }
        a[p + (size_t)q * n] = a[q + (size_t)p * n] = 0;
        for (int k = 0; k < n; ++k) {
          double vkp = V[k + (size_t)p * n], vkq = V[k + (size_t)q * n];
          V[k + (size_t)p * n] = vkp - s * (vkq + tau * vkp);
          V[k + (size_t)q * n] = vkq + s * (vkp - tau * vkq);
        }
      }
  }
  for (int i = 0; i < n; ++i) E[i] = a[i + (size_t)i * n];
  /* ascending order, as dsyev returns (selection sort keeps it simple) */
  for (int i = 0; i < n - 1; ++i) {
    int k = i;
    for (int j = i + 1; j < n; ++j) if (E[j] < E[k]) k = j;
    if (k != i) {
      double t = E[i]; E[i] = E[k]; E[k] = t;
      for (int r = 0; r < n; ++r) {
        double x = V[r + (size_t)i * n]; V[r + (size_t)i * n] = V[r + (size_t)k * n]; V[r + (size_t)k * n] = x;
      }
    }
  }
  free(a);
  return OKO_OK;
}

int oko_symeig(int n, const double* A, int lda, double* E, double* V) {
  if (eig_backend == 1) {
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) V[i + (size_t)j * n] = A[i + (size_t)j * lda];
    int lwork = -1, info = 0;
    double wq;
    p_dsyev("V", "L", &n, V, &n, E, &wq, &lwork, &info, 1, 1);
    lwork = (int)wq;
    double* work = (double*)malloc(sizeof(double) * (size_t)lwork);
    if (!work) return OKO_ENOMEM;
    p_dsyev("V", "L", &n, V, &n, E, work, &lwork, &info, 1, 1);
    free(work);
    return info == 0 ? OKO_OK : OKO_ENAN;
  }
  return jacobi_eig(n, A, lda, E, V);
}

/* C(m x n) = op(A) op(B) with op = 'N' or 'T' (column-major). */
static void gemm(char ta, char tb, int m, int n, int k, const double* A, int lda,
                 const double* B, int ldb, double* C, int ldc) {
  if (gemm_backend == 1) {
    double one = 1, zero = 0;
    char sa[2] = {ta, 0}, sb[2] = {tb, 0};
    p_dgemm(sa, sb, &m, &n, &k, &one, A, &lda, B, &ldb, &zero, C, &ldc, 1, 1);
    return;
  }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      double s = 0;
      for (int l = 0; l < k; ++l) {
        double a = (ta == 'N') ? A[i + (size_t)l * lda] : A[l + (size_t)i * lda];
        double b = (tb == 'N') ? B[l + (size_t)j * ldb] : B[j + (size_t)l * ldb];
        s += a * b;
      }
      C[i + (size_t)j * ldc] = s;
    }
}

/* ------------------------------------------------------------------------- */
/* perpSpace (matoper.F90:509-535) and RotateVector (rrsqrt.F90:737-745).      */
/* ------------------------------------------------------------------------- */
/* H: n x (n-1), column-major, ld n.  Orthonormal basis of w's complement
 * (Hoteit et al. 2002). */
void oko_perp_space(int n, const double* w, double* H) {
  double alpha = -1. / (fabs(w[n - 1]) + 1.);
  double sgn = (w[n - 1] >= 0) ? 1. : -1.; /* Fortran sign(1.,x): +1 for x >= +0 */
  for (int j = 0; j < n - 1; ++j) {
    for (int i = 0; i < n - 1; ++i) {
      double h = alpha * w[i] * w[j];
      if (i == j) h = h + 1.;
      H[i + (size_t)j * n] = h;
    }
    H[(n - 1) + (size_t)j * n] = alpha * (w[n - 1] + sgn) * w[j];
  }
}

/* Omega (n x n) = v w^T + P(v) P(w)^T, orthogonal with Omega·w = v. */
int oko_rotate_vector(int n, const double* w, const double* v, double* Omega) {
  double* Pv = (double*)malloc(sizeof(double) * (size_t)n * (n - 1));
  double* Pw = (double*)malloc(sizeof(double) * (size_t)n * (n - 1));
  if (!Pv || !Pw) { free(Pv); free(Pw); return OKO_ENOMEM; }
  oko_perp_space(n, v, Pv);
  oko_perp_space(n, w, Pw);
  gemm('N', 'T', n, n, n - 1, Pv, n, Pw, n, Omega, n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) Omega[i + (size_t)j * n] = v[i] * w[j] + Omega[i + (size_t)j * n];
  free(Pv);
  free(Pw);
  return OKO_OK;
}

/* ------------------------------------------------------------------------- */
/* analysisIncrement (rrsqrt.F90:100-190) with R = DCDCovar(w, DiagCovar(σ²)) */
/* (covariance.F90:612-619 and 425-432): R⁻¹x = w·((w·x)/σ²).                */
/* An optional exclusion mask d ∈ {0,1} wraps R once more                     */
/* (assimilation.F90:3086-3092): R⁻¹x = w·(d·((d·(w·x))/σ²)).                */
/*   p obs rows; Sf s x N (ldSf); HSf p x N (ldHS); outputs xa_xf (s),        */
/*   Sa s x N (ldSa), ampl (N, optional).                                     */
/* ------------------------------------------------------------------------- */
static inline double rinv_apply(double w, double dex, int has_ex, double sig2, double x) {
  if (has_ex) return w * (dex * ((dex * (w * x)) / sig2));
  return w * ((w * x) / sig2);
}

int oko_analysis_increment(int64_t p, int N, int64_t s, const double* Hxf, const double* yo,
                           const double* w, const double* sig2, const double* dex,
                           const double* Sf, int64_t ldSf, const double* HSf, int64_t ldHS,
                           double* xa_xf, double* Sa, int64_t ldSa, double* ampl_out) {
  size_t NN = (size_t)N * N;
  double* Q = (double*)malloc(sizeof(double) * (size_t)p * N);
  double* temp = (double*)malloc(sizeof(double) * NN);
  double* U = (double*)malloc(sizeof(double) * NN);
  double* lam = (double*)malloc(sizeof(double) * N);
  double* tmpN = (double*)malloc(sizeof(double) * N * 4);
  double* Om = (double*)malloc(sizeof(double) * NN);
  double* X1 = (double*)malloc(sizeof(double) * NN);
  double* X3 = (double*)malloc(sizeof(double) * NN);
  double* innov = (double*)malloc(sizeof(double) * (p > 0 ? p : 1));
  int rc = OKO_OK;
  if (!Q || !temp || !U || !lam || !tmpN || !Om || !X1 || !X3 || !innov) { rc = OKO_ENOMEM; goto done; }
  int has_ex = dex != 0;

  /* temp = HSf .tx. (R%mldivide(HSf))   (rrsqrt.F90:135) */
  for (int k = 0; k < N; ++k)
    for (int64_t l = 0; l < p; ++l)
      Q[l + (size_t)k * p] = rinv_apply(w[l], has_ex ? dex[l] : 1, has_ex, sig2[l], HSf[l + (size_t)k * ldHS]);
  gemm('T', 'N', N, N, (int)p, HSf, (int)ldHS, Q, (int)p, temp, N);
  /* call symeig(temp,lambda,U); where (lambda < 0) lambda = 0; lambda = (1+lambda)**(-1) */
  rc = oko_symeig(N, temp, N, lam, U);
  if (rc) goto done;
  for (int i = 0; i < N; ++i) {
    if (lam[i] < 0) lam[i] = 0;
    lam[i] = 1. / (1. + lam[i]);
  }
  /* ampl = U.x.(lambda.dx.(U.tx.(HSf.tx.(R%mldivide(yo-Hxf)))))  (rrsqrt.F90:142) */
  {
    double* b = tmpN;       /* HSf^T R^-1 (yo-Hxf) */
    double* c = tmpN + N;   /* lambda .* U^T b */
    double* ampl = tmpN + 2 * N;
    for (int64_t l = 0; l < p; ++l)
      innov[l] = rinv_apply(w[l], has_ex ? dex[l] : 1, has_ex, sig2[l], yo[l] - Hxf[l]);
    for (int k = 0; k < N; ++k) {
      double t = 0;
      for (int64_t l = 0; l < p; ++l) t += HSf[l + (size_t)k * ldHS] * innov[l];
      b[k] = t;
    }
    for (int k = 0; k < N; ++k) {
      double t = 0;
      for (int i = 0; i < N; ++i) t += U[i + (size_t)k * N] * b[i];
      c[k] = lam[k] * t;
    }
    for (int i = 0; i < N; ++i) {
      double t = 0;
      for (int k = 0; k < N; ++k) t += U[i + (size_t)k * N] * c[k];
      ampl[i] = t;
    }
    for (int i = 0; i < N; ++i)
      if (ampl[i] != ampl[i]) { rc = OKO_ENAN; goto done; } /* rrsqrt.F90:145-149 */
    /* xa_xf = Sf.x.ampl */
    for (int64_t r = 0; r < s; ++r) {
      double t = 0;
      for (int k = 0; k < N; ++k) t += Sf[r + (size_t)k * ldSf] * ampl[k];
      xa_xf[r] = t;
    }
    if (ampl_out) memcpy(ampl_out, ampl, sizeof(double) * N);
  }
  if (Sa) {
    /* ROTATE_ENSEMBLE branch (rrsqrt.F90:158-185) */
    double* sq = tmpN;      /* sqrt_lambda */
    double* v = tmpN + N;
    double* wv = tmpN + 2 * N;
    double* cs = tmpN + 3 * N;
    for (int i = 0; i < N; ++i) sq[i] = sqrt(lam[i]);
    for (int i = 0; i < N; ++i) wv[i] = 1. / sqrt(1. * N);
    for (int k = 0; k < N; ++k) {
      double t = 0;
      for (int i = 0; i < N; ++i) t += U[i + (size_t)k * N];
      cs[k] = t / sq[k];
    }
    for (int i = 0; i < N; ++i) {
      double t = 0;
      for (int k = 0; k < N; ++k) t += U[i + (size_t)k * N] * cs[k];
      v[i] = t;
    }
    double nv = 0;
    for (int i = 0; i < N; ++i) nv += v[i] * v[i];
    nv = sqrt(nv);
    for (int i = 0; i < N; ++i) v[i] = v[i] / nv;
    rc = oko_rotate_vector(N, wv, v, Om);
    if (rc) goto done;
    /* Sa = Sf.x.(U.x.(sqrt_lambda.dx.(U.tx.Omega))) */
    gemm('T', 'N', N, N, N, U, N, Om, N, X1, N);
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < N; ++i) X1[i + (size_t)j * N] *= sq[i];
    gemm('N', 'N', N, N, N, U, N, X1, N, X3, N);
    gemm('N', 'N', (int)s, N, N, Sf, (int)ldSf, X3, N, Sa, (int)ldSa);
  }
done:
  free(Q); free(temp); free(U); free(lam); free(tmpN); free(Om); free(X1); free(X3); free(innov);
  return rc;
}

/* ------------------------------------------------------------------------- */
/* locAnalysisIncrement + locAnalysis (rrsqrt.F90:258-466).                    */
/* The selection is given per zone in CSR form (callback-exact mode):          */
/*   obs sel_idx[sel_ptr[z] .. sel_ptr[z+1]) with taper weights sel_w.         */
/* localise_obs = 1 (the default, rrsqrt.F90:313): only the listed obs are     */
/* packed (rrsqrt.F90:395-404).  localise_obs = 0: the list must hold all m    */
/* obs in order (weights of the non-relevant ones included, as OAK computes    */
/* them) and amplitudes are returned (rrsqrt.F90:378-383).                     */
/* A zone with an empty list keeps xa = xf, Sa = Sf (rrsqrt.F90:322-326,371).  */
/* ------------------------------------------------------------------------- */
int oko_loc_analysis_csr(int nz, const int64_t* zone_ptr, int N, int64_t m,
                         const int64_t* sel_ptr, const int32_t* sel_idx, const double* sel_w,
                         const double* xf, const double* Hxf, const double* yo,
                         const double* sig2, const double* dex,
                         const double* Sf, int64_t ldSf, const double* HSf, int64_t ldHS,
                         double* xa, double* Sa, int64_t ldSa, double* ampl, int localise_obs,
                         int nthreads) {
  int64_t n = zone_ptr[nz];
  int rc = OKO_OK;
  for (int64_t i = 0; i < n; ++i) xa[i] = 0;
  if (Sa)
    for (int k = 0; k < N; ++k)
      for (int64_t i = 0; i < n; ++i) Sa[i + (size_t)k * ldSa] = Sf[i + (size_t)k * ldSf];
  if (ampl) memset(ampl, 0, sizeof(double) * (size_t)N * nz);
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(| : rc)
#endif
  {
    int64_t pmax = 0;
    for (int z = 0; z < nz; ++z) {
      int64_t pz = sel_ptr[z + 1] - sel_ptr[z];
      if (pz > pmax) pmax = pz;
    }
    if (pmax < 1) pmax = 1;
    double* HSz = (double*)malloc(sizeof(double) * (size_t)pmax * N);
    double* buf = (double*)malloc(sizeof(double) * (size_t)pmax * 5);
    if (!HSz || !buf) rc |= OKO_ENOMEM;
    else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
      for (int z = 0; z < nz; ++z) {
        int64_t p = sel_ptr[z + 1] - sel_ptr[z];
        if (p == 0) continue;
        const int32_t* idx = sel_idx + sel_ptr[z];
        double* yz = buf;
        double* hz = buf + pmax;
        double* wz = buf + 2 * pmax;
        double* sz = buf + 3 * pmax;
        double* dz = buf + 4 * pmax;
        for (int64_t l = 0; l < p; ++l) {
          int32_t o = idx[l];
          yz[l] = yo[o]; hz[l] = Hxf[o]; wz[l] = sel_w[sel_ptr[z] + l]; sz[l] = sig2[o];
          dz[l] = dex ? dex[o] : 1;
          for (int k = 0; k < N; ++k) HSz[l + (size_t)k * p] = HSf[o + (size_t)k * ldHS];
        }
        int64_t i1 = zone_ptr[z], s = zone_ptr[z + 1] - zone_ptr[z];
        int r = oko_analysis_increment(p, N, s, hz, yz, wz, sz, dex ? dz : 0, Sf + i1, ldSf, HSz, p,
                                       xa + i1, Sa ? Sa + i1 : 0, ldSa,
                                       (ampl && !localise_obs) ? ampl + (size_t)z * N : 0);
        if (r) rc |= r;
      }
    }
    free(HSz);
    free(buf);
  }
  for (int64_t i = 0; i < n; ++i) xa[i] = xf[i] + xa[i]; /* rrsqrt.F90:463 */
  return rc;
}

/* Geometry mode: the zone loop of locAnalysisIncrement with OAK's own
 * selectObservations (assimilation.F90:3683-3771) evaluated per zone.
 * zpos: nz x 4 (x,y,z,t of the zone's first element, row per zone);
 * L, cutoff: per zone (corrLength/maxLength of the first element).
 * Only zones [z0, z1) are analysed (the others keep xa = xf, Sa = Sf); this
 * lets the CPU baseline and the parity tests run a bounded sample. */
int oko_loc_analysis_geo(int nz, const int64_t* zone_ptr, int z0, int z1, int N, int64_t m,
                         int metric, int loctype, const double* zpos, const double* L,
                         const double* cutoff, const double* ox, const double* oy,
                         const double* oz, const double* ot,
                         const double* xf, const double* Hxf, const double* yo,
                         const double* sig2, const double* dex,
                         const double* Sf, int64_t ldSf, const double* HSf, int64_t ldHS,
                         double* xa, double* Sa, int64_t ldSa, double* ampl, int localise_obs,
                         int nthreads, int64_t* counts) {
  int64_t n = zone_ptr[nz];
  int rc = OKO_OK;
  for (int64_t i = 0; i < n; ++i) xa[i] = 0;
  if (Sa)
    for (int k = 0; k < N; ++k)
      for (int64_t i = 0; i < n; ++i) Sa[i + (size_t)k * ldSa] = Sf[i + (size_t)k * ldSf];
  if (ampl) memset(ampl, 0, sizeof(double) * (size_t)N * nz);
  if (z0 < 0) z0 = 0;
  if (z1 > nz) z1 = nz;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(| : rc)
#endif
  {
    double* weight = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    unsigned char* rel = (unsigned char*)malloc((size_t)(m > 0 ? m : 1));
    double* HSz = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1) * N);
    double* buf = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1) * 5);
    if (!weight || !rel || !HSz || !buf) rc |= OKO_ENOMEM;
    else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
      for (int z = z0; z < z1; ++z) {
        int64_t cnt = oko_select(metric, loctype, zpos + 4 * (size_t)z, L[z], cutoff[z], m, ox, oy,
                                 oz, ot, weight, rel);
        if (counts) counts[z] = cnt;
        if (cnt == 0) continue;
        double* yz = buf;
        double* hz = buf + m;
        double* wz = buf + 2 * m;
        double* sz = buf + 3 * m;
        double* dz = buf + 4 * m;
        int64_t p = 0;
        for (int64_t l = 0; l < m; ++l) {
          if (localise_obs && !rel[l]) continue;
          yz[p] = yo[l]; hz[p] = Hxf[l]; wz[p] = weight[l]; sz[p] = sig2[l];
          dz[p] = dex ? dex[l] : 1;
          ++p;
        }
        int64_t q = 0;
        for (int64_t l = 0; l < m; ++l) {
          if (localise_obs && !rel[l]) continue;
          for (int k = 0; k < N; ++k) HSz[q + (size_t)k * p] = HSf[l + (size_t)k * ldHS];
          ++q;
        }
        int64_t i1 = zone_ptr[z], s = zone_ptr[z + 1] - zone_ptr[z];
        int r = oko_analysis_increment(p, N, s, hz, yz, wz, sz, dex ? dz : 0, Sf + i1, ldSf, HSz, p,
                                       xa + i1, Sa ? Sa + i1 : 0, ldSa,
                                       (ampl && !localise_obs) ? ampl + (size_t)z * N : 0);
        if (r) rc |= r;
      }
    }
    free(weight); free(rel); free(HSz); free(buf);
  }
  for (int64_t i = 0; i < n; ++i) xa[i] = xf[i] + xa[i];
  return rc;
}

/* ------------------------------------------------------------------------- */
/* Observation operator: COO sparse matrix times vector                       */
/* (matoper_inc.F90:220-242; SparseMatrix matoper.F90:30-39).                 */
/* ------------------------------------------------------------------------- */
void oko_obsoper(int64_t m, int64_t nnz, const int32_t* Hi, const int32_t* Hj, const double* Hs,
                 const double* x, double* Hx) {
  for (int64_t l = 0; l < m; ++l) Hx[l] = 0;
  for (int64_t k = 0; k < nnz; ++k) Hx[Hi[k]] = Hx[Hi[k]] + Hs[k] * x[Hj[k]];
}

/* ------------------------------------------------------------------------- */
/* Ensemble mode, as Assim runs it with schemetype = LocalScheme:              */
/*   HE = H·E + Hshift; means; anomalies scaled by sqrt(N-1)                  */
/*   (assimilation.F90:3094-3135, ppdef.h:52); locAnalysis (geometry mode);    */
/*   inflation, maxCorrection clamp, ensemble rebuild and re-centring          */
/*   (assimilation.F90:3299-3357, 3558-3563).                                 */
/* E, Ea: n x N (ld n), in zone-permuted order; H columns already permuted.    */
/* maxCorrVec: OAK's per-element maxCorrection (assimilation.F90:462-476,      */
/* permuted like the state), or NULL to use the scalar maxCorrection for all.  */
/* Zones outside [z0,z1) are left without analysis (testing only).           */
/* ------------------------------------------------------------------------- */
int oko_ensemble_analysis(int nz, const int64_t* zone_ptr, int z0, int z1, int N, int64_t m,
                          int metric, int loctype, const double* zpos, const double* L,
                          const double* cutoff, const double* ox, const double* oy,
                          const double* oz, const double* ot,
                          int64_t nnz, const int32_t* Hi, const int32_t* Hj, const double* Hs,
                          const double* Hshift, const double* yo, const double* sig2,
                          const double* dex, const double* E, double inflation,
                          double maxCorrection, const double* maxCorrVec, double* Ea,
                          double* xa_out, int nthreads, int64_t* counts) {
  int64_t n = zone_ptr[nz];
  double scaling = sqrt((double)N - 1.);
  double* HE = (double*)malloc(sizeof(double) * (size_t)m * N);
  double* xf = (double*)malloc(sizeof(double) * (size_t)n);
  double* Hxf = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  double* Sf = (double*)malloc(sizeof(double) * (size_t)n * N);
  double* xa = (double*)malloc(sizeof(double) * (size_t)n);
  int rc = OKO_OK;
  if (!HE || !xf || !Hxf || !Sf || !xa) { rc = OKO_ENOMEM; goto done; }
  for (int k = 0; k < N; ++k) {
    oko_obsoper(m, nnz, Hi, Hj, Hs, E + (size_t)k * n, HE + (size_t)k * m);
    for (int64_t l = 0; l < m; ++l) HE[l + (size_t)k * m] = HE[l + (size_t)k * m] + Hshift[l];
  }
  for (int64_t i = 0; i < n; ++i) {
    double t = 0;
    for (int k = 0; k < N; ++k) t += E[i + (size_t)k * n];
    xf[i] = t / N;
  }
  for (int64_t l = 0; l < m; ++l) {
    double t = 0;
    for (int k = 0; k < N; ++k) t += HE[l + (size_t)k * m];
    Hxf[l] = t / N;
  }
  for (int k = 0; k < N; ++k) {
    for (int64_t i = 0; i < n; ++i) Sf[i + (size_t)k * n] = (E[i + (size_t)k * n] - xf[i]) / scaling;
    for (int64_t l = 0; l < m; ++l) HE[l + (size_t)k * m] = (HE[l + (size_t)k * m] - Hxf[l]) / scaling;
  }
  rc = oko_loc_analysis_geo(nz, zone_ptr, z0, z1, N, m, metric, loctype, zpos, L, cutoff, ox, oy,
                            oz, ot, xf, Hxf, yo, sig2, dex, Sf, n, HE, m, xa, Ea, n, 0, 1,
                            nthreads, counts);
  if (rc) goto done;
  /* Ea currently holds Sa (error modes) */
  if (inflation != 1)
    for (size_t i = 0; i < (size_t)n * N; ++i) Ea[i] = inflation * Ea[i];
  for (int64_t i = 0; i < n; ++i) {
    const double mc = maxCorrVec ? maxCorrVec[i] : maxCorrection;
    if (xa[i] - mc > xf[i]) xa[i] = xf[i] + mc;
    if (xa[i] < xf[i] - mc) xa[i] = xf[i] - mc;
  }
  for (int k = 0; k < N; ++k)
    for (int64_t i = 0; i < n; ++i) Ea[i + (size_t)k * n] = xa[i] + Ea[i + (size_t)k * n] * scaling;
  /* re-centre: xa = mean(Ea); Sa = (Ea-xa)/scaling; Ea = xa + Sa*sqrt(N-1) */
  for (int64_t i = 0; i < n; ++i) {
    double t = 0;
    for (int k = 0; k < N; ++k) t += Ea[i + (size_t)k * n];
    xa[i] = t / N;
  }
  for (int k = 0; k < N; ++k)
    for (int64_t i = 0; i < n; ++i) {
      double sa = (Ea[i + (size_t)k * n] - xa[i]) / scaling;
      Ea[i + (size_t)k * n] = xa[i] + sa * sqrt((double)N - 1.);
    }
  if (xa_out) memcpy(xa_out, xa, sizeof(double) * (size_t)n);
done:
  free(HE); free(xf); free(Hxf); free(Sf); free(xa);
  return rc;
}

/* ------------------------------------------------------------------------- */
/* report (assimilation.F90:3584-3628): innovation statistics written to the  */
/* log for the forecast and the analysis (:3389-3391) and per observed        */
/* variable (:3410-3430).  d = yo - Hx (m), HS m x N (ldHS), sig2 the diagonal */
/* of R; dex (may be NULL) is the exclusion mask D (0 = excluded), which wraps */
/* R as DCDCovar (:3086-3092) so that R^-1 x = D (D x) / sig2                 */
/* (covariance.F90:612-619, 425-432).  With ingrid = count(D != 0):           */
/*   out[0] ingrid                                                            */
/*   out[1] rms_yo-Hx                 sqrt(sum(d^2, D != 0) / ingrid)  (:3607) */
/*   out[2] bias_yo-Hx                sum(d, D != 0) / ingrid          (:3608) */
/*   out[3] rms_invsqrtR_yo-Hx        sqrt(sum(d * R^-1 d) / ingrid)   (:3617) */
/*   out[4] projection_yo-Hx_into_HSf sqrt(sum((HS^T R^-1 d)^2)/ingrid) (:3620) */
/* ------------------------------------------------------------------------- */
void oko_report(int64_t m, int N, const double* d, const double* HS, int64_t ldHS,
                const double* sig2, const double* dex, double* out) {
  int64_t ingrid = 0;
  double s2 = 0, s1 = 0, sr = 0, pp = 0;
  for (int64_t l = 0; l < m; ++l) {
    const int valid = !dex || dex[l] != 0;
    if (valid) {
      ingrid++;
      s2 = s2 + d[l] * d[l];
      s1 = s1 + d[l];
    }
  }
  double* r = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  if (!r) { for (int i = 0; i < 5; ++i) out[i] = NAN; return; }
  for (int64_t l = 0; l < m; ++l) {
    r[l] = dex ? dex[l] * ((dex[l] * d[l]) / sig2[l]) : d[l] / sig2[l];
    sr = sr + d[l] * r[l];
  }
  for (int k = 0; k < N; ++k) {
    double t = 0;
    for (int64_t l = 0; l < m; ++l) t = t + HS[l + (size_t)k * ldHS] * r[l];
    pp = pp + t * t;
  }
  free(r);
  const double g = (double)ingrid;
  out[0] = g;
  out[1] = sqrt(s2 / g);
  out[2] = s1 / g;
  out[3] = sqrt(sr / g);
  out[4] = sqrt(pp / g);
}

/* The forecast or analysis report of Assim's ensemble path for an ensemble E  */
/* (n x N, ld n): HE = H E + Hshift, Hx = mean(HE, 2), HS = (HE - Hx)/sqrt(N-1) */
/* (assimilation.F90:3112-3134 for Ef, :3325-3357 for Ea), d = yo - Hx; then   */
/* oko_report over each observation range [range_ptr[r], range_ptr[r+1])       */
/* (R%sub and the row slices of :3424-3429); out is 5 x nranges.              */
int oko_ensemble_report(int64_t n, int N, int64_t m, int64_t nnz, const int32_t* Hi,
                        const int32_t* Hj, const double* Hs, const double* Hshift,
                        const double* yo, const double* sig2, const double* dex, const double* E,
                        int nranges, const int64_t* range_ptr, double* out) {
  const double scaling = sqrt((double)N - 1.);
  double* HE = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1) * N);
  double* Hx = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  double* d = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  if (!HE || !Hx || !d) { free(HE); free(Hx); free(d); return OKO_ENOMEM; }
  for (int k = 0; k < N; ++k) {
    oko_obsoper(m, nnz, Hi, Hj, Hs, E + (size_t)k * n, HE + (size_t)k * m);
    for (int64_t l = 0; l < m; ++l) HE[l + (size_t)k * m] = HE[l + (size_t)k * m] + Hshift[l];
  }
  for (int64_t l = 0; l < m; ++l) {
    double t = 0;
    for (int k = 0; k < N; ++k) t += HE[l + (size_t)k * m];
    Hx[l] = t / N;
    d[l] = yo[l] - Hx[l];
  }
  for (int k = 0; k < N; ++k)
    for (int64_t l = 0; l < m; ++l) HE[l + (size_t)k * m] = (HE[l + (size_t)k * m] - Hx[l]) / scaling;
  for (int r = 0; r < nranges; ++r) {
    const int64_t a = range_ptr[r], b = range_ptr[r + 1];
    oko_report(b - a, N, d + a, HE + a, m, sig2 + a, dex ? dex + a : 0, out + 5 * r);
  }
  free(HE); free(Hx); free(d);
  return OKO_OK;
}

static int cmp_i64(const void* a, const void* b) {
  int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  return (x > y) - (x < y);
}

/* assimilation.F90:578-641 (initPartition) after the gap relabelling of
 * :408-412.  partition: n zone labels (any integers); outputs zone_size (nz),
 * zone_index (permuted position -> original index, stable), inv_zone_index.
 * Returns the number of zones. */
int oko_partition(int64_t n, const int64_t* partition, int64_t* zone_size, int64_t* zone_index,
                  int64_t* inv_zone_index) {
  /* relabel through the sorted unique values (unique(...,ind2=partition)) */
  int64_t* sorted = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  int64_t* lab = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  if (!sorted || !lab) { free(sorted); free(lab); return -1; }
  memcpy(sorted, partition, sizeof(int64_t) * (size_t)n);
  qsort(sorted, (size_t)n, sizeof(int64_t), cmp_i64);
  int64_t nu = 0;
  for (int64_t i = 0; i < n; ++i)
    if (i == 0 || sorted[i] != sorted[i - 1]) sorted[nu++] = sorted[i];
  for (int64_t i = 0; i < n; ++i) {
    int64_t lo = 0, hi = nu - 1;
    while (lo < hi) {
      int64_t mid = (lo + hi) / 2;
      if (sorted[mid] < partition[i]) lo = mid + 1; else hi = mid;
    }
    lab[i] = lo;
  }
  for (int64_t z = 0; z < nu; ++z) zone_size[z] = 0;
  for (int64_t i = 0; i < n; ++i) zone_size[lab[i]]++;
  int64_t* c = sorted; /* reuse as cIndex */
  int64_t acc = 0;
  for (int64_t z = 0; z < nu; ++z) { c[z] = acc; acc += zone_size[z]; }
  for (int64_t i = 0; i < n; ++i) {
    int64_t z = lab[i];
    zone_index[c[z]] = i;
    inv_zone_index[i] = c[z];
    c[z]++;
  }
  free(sorted);
  free(lab);
  return (int)nu;
}
