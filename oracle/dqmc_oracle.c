/*
 * dqmc_oracle.c -- CPU FP64 restatement of MonteCarlo.jl's DQMC local-update
 * sweep.  TEST INFRASTRUCTURE ONLY: this file is the parity oracle and the
 * timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (libdqmc_hip.so) never does.
 *
 * Every routine follows a reference file (paths relative to the
 * MonteCarlo.jl repository root):
 *   stack buffers / chunking      src/flavors/DQMC/stack.jl:154-206
 *   slice-matrix multiplies       src/flavors/DQMC/stack.jl:319-367
 *   add_slice_sequence_left/right src/flavors/DQMC/stack.jl:377-416
 *   build / reverse_build_stack   src/flavors/DQMC/stack.jl:257-308
 *   calculate_greens_AVX!         src/flavors/DQMC/stack.jl:442-496
 *   calculate_greens(mc, slice)   src/flavors/DQMC/stack.jl:525-583
 *   wrap_greens! / propagate      src/flavors/DQMC/stack.jl:594-730
 *   udt_AVX_pivot! and helpers    src/flavors/DQMC/linalg/UDT.jl:157-334
 *   rdivp! / rvadd! / vinv!       src/flavors/DQMC/linalg/real.jl:118-226
 *   Hirsch fields (propose)       src/flavors/DQMC/fields.jl:363-449
 *   detratio / Sherman-Morrison   src/flavors/DQMC/fields.jl:58-84,176-286
 *   rank-1 helpers                src/flavors/DQMC/linalg/updates.jl:7-131
 *   sweep_spatial / local_sweep   src/flavors/DQMC/updates/local_updates.jl:7-60
 *   physical greens transform     src/flavors/DQMC/greens.jl:114-125
 *
 * Layout: every matrix is n x n column-major FP64; a block-diagonal matrix
 * (MagneticHirschField, 2 flavours) is `flv` consecutive n x n blocks.
 * Slice and stack indices are 1-based as in the reference.
 *
 * Parity status of each part is recorded by the tests that pin it
 * (tests/test_oracle_*.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* GEMM: C = A * B, n x n column-major.  Register-blocked kernels for AVX2   */
/* and AVX-512 chosen at run time; the reference's vmul! loops are           */
/* LoopVectorization-tiled (real.jl:7-15), so a vectorised, register-tiled   */
/* kernel is the faithful CPU cost model.                                    */
/* ------------------------------------------------------------------------ */
typedef double v4d __attribute__((vector_size(32), aligned(8)));
typedef double v8d __attribute__((vector_size(64), aligned(8)));

__attribute__((target("avx2,fma")))
static void gemm_nn_avx2(int n, const double *A, const double *B, double *C) {
    int j0 = 0;
    for (; j0 + 4 <= n; j0 += 4) {
        const double *b0 = B + (size_t)(j0 + 0) * n, *b1 = B + (size_t)(j0 + 1) * n;
        const double *b2 = B + (size_t)(j0 + 2) * n, *b3 = B + (size_t)(j0 + 3) * n;
        int i0 = 0;
        for (; i0 + 8 <= n; i0 += 8) {
            v4d c00 = {0}, c01 = {0}, c10 = {0}, c11 = {0};
            v4d c20 = {0}, c21 = {0}, c30 = {0}, c31 = {0};
            const double *a = A + i0;
            for (int k = 0; k < n; k++) {
                v4d a0 = *(const v4d *)(a + (size_t)k * n);
                v4d a1 = *(const v4d *)(a + (size_t)k * n + 4);
                v4d s0 = {b0[k], b0[k], b0[k], b0[k]};
                v4d s1 = {b1[k], b1[k], b1[k], b1[k]};
                v4d s2 = {b2[k], b2[k], b2[k], b2[k]};
                v4d s3 = {b3[k], b3[k], b3[k], b3[k]};
                c00 += a0 * s0; c01 += a1 * s0;
                c10 += a0 * s1; c11 += a1 * s1;
                c20 += a0 * s2; c21 += a1 * s2;
                c30 += a0 * s3; c31 += a1 * s3;
            }
            double *c = C + i0;
            *(v4d *)(c + (size_t)(j0 + 0) * n) = c00; *(v4d *)(c + (size_t)(j0 + 0) * n + 4) = c01;
            *(v4d *)(c + (size_t)(j0 + 1) * n) = c10; *(v4d *)(c + (size_t)(j0 + 1) * n + 4) = c11;
            *(v4d *)(c + (size_t)(j0 + 2) * n) = c20; *(v4d *)(c + (size_t)(j0 + 2) * n + 4) = c21;
            *(v4d *)(c + (size_t)(j0 + 3) * n) = c30; *(v4d *)(c + (size_t)(j0 + 3) * n + 4) = c31;
        }
        for (; i0 < n; i0++) {
            double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
            for (int k = 0; k < n; k++) {
                double a = A[i0 + (size_t)k * n];
                s0 += a * b0[k]; s1 += a * b1[k]; s2 += a * b2[k]; s3 += a * b3[k];
            }
            C[i0 + (size_t)(j0 + 0) * n] = s0; C[i0 + (size_t)(j0 + 1) * n] = s1;
            C[i0 + (size_t)(j0 + 2) * n] = s2; C[i0 + (size_t)(j0 + 3) * n] = s3;
        }
    }
    for (; j0 < n; j0++) {
        for (int i = 0; i < n; i++) C[i + (size_t)j0 * n] = 0.0;
        for (int k = 0; k < n; k++) {
            double b = B[k + (size_t)j0 * n];
            for (int i = 0; i < n; i++) C[i + (size_t)j0 * n] += A[i + (size_t)k * n] * b;
        }
    }
}

__attribute__((target("avx512f,fma")))
static void gemm_nn_avx512(int n, const double *A, const double *B, double *C) {
    int j0 = 0;
    for (; j0 + 4 <= n; j0 += 4) {
        const double *b0 = B + (size_t)(j0 + 0) * n, *b1 = B + (size_t)(j0 + 1) * n;
        const double *b2 = B + (size_t)(j0 + 2) * n, *b3 = B + (size_t)(j0 + 3) * n;
        int i0 = 0;
        for (; i0 + 16 <= n; i0 += 16) {
            v8d c00 = {0}, c01 = {0}, c10 = {0}, c11 = {0};
            v8d c20 = {0}, c21 = {0}, c30 = {0}, c31 = {0};
            const double *a = A + i0;
            for (int k = 0; k < n; k++) {
                v8d a0 = *(const v8d *)(a + (size_t)k * n);
                v8d a1 = *(const v8d *)(a + (size_t)k * n + 8);
                double x0 = b0[k], x1 = b1[k], x2 = b2[k], x3 = b3[k];
                v8d s0 = {x0, x0, x0, x0, x0, x0, x0, x0};
                v8d s1 = {x1, x1, x1, x1, x1, x1, x1, x1};
                v8d s2 = {x2, x2, x2, x2, x2, x2, x2, x2};
                v8d s3 = {x3, x3, x3, x3, x3, x3, x3, x3};
                c00 += a0 * s0; c01 += a1 * s0;
                c10 += a0 * s1; c11 += a1 * s1;
                c20 += a0 * s2; c21 += a1 * s2;
                c30 += a0 * s3; c31 += a1 * s3;
            }
            double *c = C + i0;
            *(v8d *)(c + (size_t)(j0 + 0) * n) = c00; *(v8d *)(c + (size_t)(j0 + 0) * n + 8) = c01;
            *(v8d *)(c + (size_t)(j0 + 1) * n) = c10; *(v8d *)(c + (size_t)(j0 + 1) * n + 8) = c11;
            *(v8d *)(c + (size_t)(j0 + 2) * n) = c20; *(v8d *)(c + (size_t)(j0 + 2) * n + 8) = c21;
            *(v8d *)(c + (size_t)(j0 + 3) * n) = c30; *(v8d *)(c + (size_t)(j0 + 3) * n + 8) = c31;
        }
        for (; i0 < n; i0++) {
            double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
            for (int k = 0; k < n; k++) {
                double a = A[i0 + (size_t)k * n];
                s0 += a * b0[k]; s1 += a * b1[k]; s2 += a * b2[k]; s3 += a * b3[k];
            }
            C[i0 + (size_t)(j0 + 0) * n] = s0; C[i0 + (size_t)(j0 + 1) * n] = s1;
            C[i0 + (size_t)(j0 + 2) * n] = s2; C[i0 + (size_t)(j0 + 3) * n] = s3;
        }
    }
    for (; j0 < n; j0++) {
        for (int i = 0; i < n; i++) C[i + (size_t)j0 * n] = 0.0;
        for (int k = 0; k < n; k++) {
            double b = B[k + (size_t)j0 * n];
            for (int i = 0; i < n; i++) C[i + (size_t)j0 * n] += A[i + (size_t)k * n] * b;
        }
    }
}

static int g_have_avx512 = -1;

static void gemm_nn(int n, const double *A, const double *B, double *C) {
    if (g_have_avx512 < 0) {
        __builtin_cpu_init();
        g_have_avx512 = __builtin_cpu_supports("avx512f") ? 1 : 0;
    }
    if (g_have_avx512) gemm_nn_avx512(n, A, B, C);
    else gemm_nn_avx2(n, A, B, C);
}

static void transpose(int n, const double *A, double *At) {
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) At[j + (size_t)i * n] = A[i + (size_t)j * n];
}

/* C = op(A) * op(B); `work` holds n*n doubles (2 n*n when both transposed). */
static void gemm(int n, int ta, const double *A, int tb, const double *B, double *C, double *work) {
    const double *a = A, *b = B;
    if (ta) { transpose(n, A, work); a = work; }
    if (tb) { double *w = work + (ta ? (size_t)n * n : 0); transpose(n, B, w); b = w; }
    gemm_nn(n, a, b, C);
}

/* C = A * diag(d)        (real.jl:16-20) */
static void mul_diag_right(int n, double *C, const double *A, const double *d) {
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) C[i + (size_t)j * n] = A[i + (size_t)j * n] * d[j];
}
/* C = diag(d) * A        (real.jl:27-31) */
static void mul_diag_left(int n, double *C, const double *d, const double *A) {
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) C[i + (size_t)j * n] = d[i] * A[i + (size_t)j * n];
}
static void set_identity(int n, double *A) {
    memset(A, 0, sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; i++) A[i + (size_t)i * n] = 1.0;
}

/* ------------------------------------------------------------------------ */
/* Pivoted UDT, UDT.jl:157-334                                              */
/* ------------------------------------------------------------------------ */
#define M_(A, i, j) (A)[(size_t)(i) + (size_t)(j) * n]

/* indmaxcolumn, UDT.jl:175-192 (0-based j) */
static int indmaxcolumn(int n, const double *A, int j, double *maxval) {
    double squared_norm = 0.0;
    for (int k = j; k < n; k++) squared_norm += M_(A, k, j) * M_(A, k, j);
    int ii = j;
    for (int i = j + 1; i < n; i++) {
        double mi = 0.0;
        for (int k = j; k < n; k++) mi += M_(A, k, i) * M_(A, k, i);
        if (fabs(mi) > squared_norm) { squared_norm = mi; ii = i; }
    }
    *maxval = squared_norm;
    return ii;
}

/* reflector!(x, normu, j, n), UDT.jl:157-172 */
static double reflector(int n, double *x, double normu, int j) {
    double xi1 = M_(x, j, j);
    if (normu == 0.0) return 0.0;
    normu = sqrt(normu);
    double nu = copysign(normu, xi1);
    xi1 += nu;
    M_(x, j, j) = -nu;
    for (int i = j + 1; i < n; i++) M_(x, i, j) /= xi1;
    return xi1 / nu;
}

/* reflectorApply!(M, tau, k, n), UDT.jl:53-70 */
static void reflector_apply(int n, double *Mx, double tau, int k) {
    for (int j = k + 1; j < n; j++) {
        double vAj = M_(Mx, k, j);
        for (int i = k + 1; i < n; i++) vAj += M_(Mx, i, k) * M_(Mx, i, j);
        vAj = tau * vAj;
        M_(Mx, k, j) -= vAj;
        for (int i = k + 1; i < n; i++) M_(Mx, i, j) -= M_(Mx, i, k) * vAj;
    }
}

/* udt_AVX_pivot!(U, D, input, pivot, temp, apply_pivot), UDT.jl:216-334.
 * pivot is 0-based here (reference 1-based). */
static void udt_pivot(int n, double *U, double *D, double *input, int *pivot, double *temp,
                      int apply_pivot) {
    for (int i = 0; i < n; i++) pivot[i] = i;
    for (int j = 0; j < n; j++) {
        double maxval;
        int jm = indmaxcolumn(n, input, j, &maxval);
        if (jm != j) {
            int tp = pivot[jm]; pivot[jm] = pivot[j]; pivot[j] = tp;
            for (int i = 0; i < n; i++) {
                double t = M_(input, i, jm); M_(input, i, jm) = M_(input, i, j); M_(input, i, j) = t;
            }
        }
        double tau = reflector(n, input, maxval, j);
        temp[j] = tau;
        reflector_apply(n, input, tau, j);
    }
    /* explicit Q by backward accumulation, UDT.jl:271-290 */
    set_identity(n, U);
    M_(U, n - 1, n - 1) -= temp[n - 1];
    for (int k = n - 2; k >= 0; k--) {
        for (int j = k; j < n; j++) {
            double vBj = M_(U, k, j);
            for (int i = k + 1; i < n; i++) vBj += M_(input, i, k) * M_(U, i, j);
            vBj = temp[k] * vBj;
            M_(U, k, j) -= vBj;
            for (int i = k + 1; i < n; i++) M_(U, i, j) -= M_(input, i, k) * vBj;
        }
    }
    /* D = |diag R|, zero -> 1, UDT.jl:292-302 */
    for (int i = 0; i < n; i++) {
        double x = fabs(M_(input, i, i));
        D[i] = (x == 0.0) ? 1.0 : x;
    }
    /* _apply_pivot!, UDT.jl:311-334 */
    if (apply_pivot) {
        for (int i = 0; i < n; i++) {
            double d = 1.0 / D[i];
            for (int j = 0; j < i; j++) temp[pivot[j]] = 0.0;
            for (int j = i; j < n; j++) temp[pivot[j]] = d * M_(input, i, j);
            for (int j = 0; j < n; j++) M_(input, i, j) = temp[j];
        }
    } else {
        for (int i = 0; i < n; i++) {
            double d = 1.0 / D[i];
            for (int j = i; j < n; j++) M_(input, i, j) = d * M_(input, i, j);
        }
    }
}

/* rdivp!(A, T, O, pivot): A <- A[:, pivot] * triu(T)^-1, real.jl:198-226 */
static void rdivp(int n, double *A, const double *T, double *O, const int *pivot) {
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) M_(O, i, j) = M_(A, i, pivot[j]);
    for (int i = 0; i < n; i++) M_(A, i, 0) = M_(O, i, 0) / M_(T, 0, 0);
    for (int j = 1; j < n; j++) {
        for (int i = 0; i < n; i++) {
            double x = M_(O, i, j);
            for (int k = 0; k < j; k++) x -= M_(A, i, k) * M_(T, k, j);
            M_(A, i, j) = x / M_(T, j, j);
        }
    }
}

/* calculate_greens_AVX!(Ul, Dl, Tl, Ur, Dr, Tr, G, pivot, temp), stack.jl:442-496.
 * One block; every input is overwritten.  work: 2 n*n doubles. */
static void calc_greens_udt(int n, double *Ul, double *Dl, double *Tl, double *Ur, double *Dr,
                            double *Tr, double *G, int *pivot, double *temp, double *work) {
    gemm(n, 0, Tl, 1, Tr, G, work);                 /* G  = Tl * Tr'            */
    mul_diag_right(n, Tr, G, Dr);                   /* Tr = G * Dr              */
    mul_diag_left(n, G, Dl, Tr);                    /* G  = Dl * Tr             */
    udt_pivot(n, Tr, Dr, G, pivot, temp, 0);        /* G -> Tr Dr (G P)         */
    gemm(n, 0, Ul, 0, Tr, Tl, work);                /* Tl = Ul * Tr             */
    rdivp(n, Ur, G, Ul, pivot);                     /* Ur = Ur (G P)^-1         */
    gemm(n, 1, Tl, 0, Ur, Tr, work);                /* Tr = Tl' * Ur            */
    for (int i = 0; i < n; i++) M_(Tr, i, i) += Dr[i]; /* rvadd!                */
    udt_pivot(n, Ul, Dr, Tr, pivot, temp, 0);       /* Tr -> Ul Dr (Tr P)       */
    rdivp(n, Ur, Tr, G, pivot);                     /* Ur = Ur (Tr P)^-1        */
    gemm(n, 0, Tl, 0, Ul, Tr, work);                /* Tr = Tl * Ul             */
    for (int i = 0; i < n; i++) Dl[i] = 1.0 / Dr[i]; /* vinv!                   */
    mul_diag_right(n, Ul, Ur, Dl);                  /* Ul = Ur * Dl             */
    gemm(n, 0, Ul, 1, Tr, G, work);                 /* G  = Ul * Tr'            */
}

/* ------------------------------------------------------------------------ */
/* Symmetric eigen-decomposition (cyclic Jacobi) for exp(c T)               */
/* The reference uses LinearAlgebra.exp on the Hermitian T (stack.jl:235-239)*/
/* ------------------------------------------------------------------------ */
static void jacobi_eigen(int n, double *A, double *V, double *w) {
    set_identity(n, V);
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < j; i++) off += M_(A, i, j) * M_(A, i, j);
        if (off < 1e-300) break;
        for (int p = 0; p < n - 1; p++) {
            for (int q = p + 1; q < n; q++) {
                double apq = M_(A, p, q);
                if (fabs(apq) < 1e-300) continue;
                double app = M_(A, p, p), aqq = M_(A, q, q);
                double theta = (aqq - app) / (2.0 * apq);
                double t = copysign(1.0, theta) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = M_(A, k, p), akq = M_(A, k, q);
                    M_(A, k, p) = c * akp - s * akq;
                    M_(A, k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = M_(A, p, k), aqk = M_(A, q, k);
                    M_(A, p, k) = c * apk - s * aqk;
                    M_(A, q, k) = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = M_(V, k, p), vkq = M_(V, k, q);
                    M_(V, k, p) = c * vkp - s * vkq;
                    M_(V, k, q) = s * vkp + c * vkq;
                }
            }
        }
    }
    for (int i = 0; i < n; i++) w[i] = M_(A, i, i);
}

/* out = exp(c * T) for symmetric T, symmetrised like LinearAlgebra's
 * Hermitian exp. */
EXPORT void orc_expm_sym(int n, const double *T, double c, double *out) {
    double *A = malloc(sizeof(double) * (size_t)n * n);
    double *V = malloc(sizeof(double) * (size_t)n * n);
    double *W = malloc(sizeof(double) * (size_t)n * n);
    double *w = malloc(sizeof(double) * (size_t)n);
    memcpy(A, T, sizeof(double) * (size_t)n * n);
    jacobi_eigen(n, A, V, w);
    for (int j = 0; j < n; j++) {
        double e = exp(c * w[j]);
        for (int i = 0; i < n; i++) M_(W, i, j) = M_(V, i, j) * e;
    }
    /* out = W * V' */
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            double s = 0.0;
            for (int k = 0; k < n; k++) s += M_(W, i, k) * M_(V, j, k);
            M_(out, i, j) = s;
        }
    for (int j = 0; j < n; j++)
        for (int i = 0; i < j; i++) {
            double s = 0.5 * (M_(out, i, j) + M_(out, j, i));
            M_(out, i, j) = s; M_(out, j, i) = s;
        }
    free(A); free(V); free(W); free(w);
}

/* ------------------------------------------------------------------------ */
/* Uniform stream: supplied array, or SplitMix64 counter stream             */
/* ------------------------------------------------------------------------ */
static inline uint64_t splitmix64_at(uint64_t seed, uint64_t k) {
    uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
EXPORT double orc_uniform_at(uint64_t seed, uint64_t k) {
    return (double)(splitmix64_at(seed, k) >> 11) * 0x1.0p-53;
}

/* ------------------------------------------------------------------------ */
/* Simulation state (DQMCStack, stack.jl:1-74)                              */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n, flv, M, safe_mult, K;
    int check_propagation_error;
    double dtau, alpha;
    int *rstart, *rstop; /* chunk ranges, 1-based inclusive (generate_chunks) */
    /* hopping exponentials, one n x n (same for every flavour block) */
    double *eTmh, *eTph, *eT2m, *eT2p;
    /* field tables, host computed once */
    double ev_tab[2];    /* exp(-alpha), exp(+alpha)                        */
    double d_first[2];   /* Delta of block 1 for x = -1 / +1                 */
    double d_second[2];  /* Delta of block 2 (magnetic) for x = -1 / +1      */
    double bweight[2];   /* exp(-dE_boson) (density) for x = -1 / +1          */
    int8_t *conf;        /* n x M column-major */
    double *u_stack, *d_stack, *t_stack; /* (K+1) slots x flv blocks */
    double *greens, *greens_temp, *Ul, *Ur, *Tl, *Tr, *Dl, *Dr;
    double *tmp1, *tmp2, *curr_U, *eV, *work;
    int *pivot;
    double *tempv, *IG, *Gr;
    int current_slice, current_range, direction;
    /* randomness */
    const double *uniforms;
    int64_t n_uniforms;
    uint64_t seed;
    int64_t u_pos;
    int stream_exhausted;
    /* statistics */
    int64_t accepted, proposed, negative_prob, ties;
    int64_t prop_err_count;
    double prop_err_max;
    double *prop_err_log; int64_t prop_err_cap;
    double prop_err_last; int64_t stabilizations;
    /* optional per-proposal logs */
    uint8_t *dec_log; double *p_log; int64_t log_cap, log_len;
} oracle_t;

#define NN ((size_t)o->n * o->n)
#define BLK(ptr, b) ((ptr) + (size_t)(b) * NN)
#define USTK(s, b) (o->u_stack + ((size_t)((s) - 1) * o->flv + (b)) * NN)
#define TSTK(s, b) (o->t_stack + ((size_t)((s) - 1) * o->flv + (b)) * NN)
#define DSTK(s, b) (o->d_stack + ((size_t)((s) - 1) * o->flv + (b)) * o->n)

static void field_tables(oracle_t *o) {
    double a = o->alpha;
    o->ev_tab[0] = exp(-a);
    o->ev_tab[1] = exp(a);
    /* Density / block 1: dE = -2 a x, Delta = exp(dE) - 1 (fields.jl:389-390) */
    o->d_first[0] = exp(-2.0 * a * -1.0) - 1.0;
    o->d_first[1] = exp(-2.0 * a * 1.0) - 1.0;
    /* Magnetic block 2: Delta = exp(-dE) - 1 (fields.jl:441-443) */
    o->d_second[0] = exp(-(-2.0 * a * -1.0)) - 1.0;
    o->d_second[1] = exp(-(-2.0 * a * 1.0)) - 1.0;
    /* Density boson factor exp(-dE) (local_updates.jl:31) */
    o->bweight[0] = exp(-(-2.0 * a * -1.0));
    o->bweight[1] = exp(-(-2.0 * a * 1.0));
}

/* generate_chunks(length, max_chunk_size), stack.jl:154-158 */
EXPORT int orc_generate_chunks(int length, int max_chunk, int *start, int *stop) {
    int nch = (length + max_chunk - 1) / max_chunk;
    double step = (double)length / nch;
    for (int i = 1; i <= nch; i++) {
        /* Julia round(Int, x): round half to even */
        start[i - 1] = (int)nearbyint((i - 1) * step) + 1;
        stop[i - 1] = (int)nearbyint(i * step);
    }
    return nch;
}

EXPORT oracle_t *orc_create(int n, int flv, int M, int safe_mult, double dtau, double alpha,
                            int check_propagation_error) {
    oracle_t *o = calloc(1, sizeof(oracle_t));
    o->n = n; o->flv = flv; o->M = M; o->safe_mult = safe_mult;
    o->dtau = dtau; o->alpha = alpha;
    o->check_propagation_error = check_propagation_error;
    o->rstart = malloc(sizeof(int) * (M + 1));
    o->rstop = malloc(sizeof(int) * (M + 1));
    o->K = orc_generate_chunks(M, safe_mult, o->rstart, o->rstop);
    size_t nn = (size_t)n * n;
    o->eTmh = calloc(nn, sizeof(double)); o->eTph = calloc(nn, sizeof(double));
    o->eT2m = calloc(nn, sizeof(double)); o->eT2p = calloc(nn, sizeof(double));
    o->conf = calloc((size_t)n * M, 1);
    for (size_t i = 0; i < (size_t)n * M; i++) o->conf[i] = 1;
    size_t slots = (size_t)(o->K + 1) * flv;
    o->u_stack = calloc(slots * nn, sizeof(double));
    o->t_stack = calloc(slots * nn, sizeof(double));
    o->d_stack = calloc(slots * n, sizeof(double));
    double **mats[] = {&o->greens, &o->greens_temp, &o->Ul, &o->Ur, &o->Tl, &o->Tr,
                       &o->tmp1, &o->tmp2, &o->curr_U};
    for (unsigned i = 0; i < sizeof(mats) / sizeof(mats[0]); i++) *mats[i] = calloc(flv * nn, sizeof(double));
    o->Dl = calloc((size_t)flv * n, sizeof(double));
    o->Dr = calloc((size_t)flv * n, sizeof(double));
    o->eV = calloc((size_t)flv * n, sizeof(double));
    o->work = calloc(2 * nn, sizeof(double));
    o->pivot = calloc((size_t)n, sizeof(int));
    o->tempv = calloc((size_t)n, sizeof(double));
    o->IG = calloc((size_t)flv * n, sizeof(double));
    o->Gr = calloc((size_t)flv * n, sizeof(double));
    o->prop_err_cap = 4096;
    o->prop_err_log = calloc(o->prop_err_cap, sizeof(double));
    field_tables(o);
    for (int b = 0; b < flv; b++) {
        set_identity(n, BLK(o->Ul, b)); set_identity(n, BLK(o->Ur, b));
        set_identity(n, BLK(o->Tl, b)); set_identity(n, BLK(o->Tr, b));
    }
    for (int i = 0; i < flv * n; i++) { o->Dl[i] = 1.0; o->Dr[i] = 1.0; }
    o->direction = 1;
    return o;
}

EXPORT void orc_destroy(oracle_t *o) {
    if (!o) return;
    free(o->rstart); free(o->rstop);
    free(o->eTmh); free(o->eTph); free(o->eT2m); free(o->eT2p); free(o->conf);
    free(o->u_stack); free(o->t_stack); free(o->d_stack);
    free(o->greens); free(o->greens_temp); free(o->Ul); free(o->Ur); free(o->Tl); free(o->Tr);
    free(o->tmp1); free(o->tmp2); free(o->curr_U); free(o->Dl); free(o->Dr); free(o->eV);
    free(o->work); free(o->pivot); free(o->tempv); free(o->IG); free(o->Gr);
    free(o->prop_err_log);
    free(o);
}

EXPORT int orc_num_chunks(oracle_t *o) { return o->K; }
EXPORT void orc_chunk(oracle_t *o, int i, int *start, int *stop) { *start = o->rstart[i]; *stop = o->rstop[i]; }

/* init_hopping_matrices, stack.jl:209-249: exponentials of -+dtau/2 T and -+dtau T */
EXPORT void orc_set_hopping(oracle_t *o, const double *T) {
    int n = o->n;
    orc_expm_sym(n, T, -0.5 * o->dtau, o->eTmh);
    orc_expm_sym(n, T, +0.5 * o->dtau, o->eTph);
    orc_expm_sym(n, T, -o->dtau, o->eT2m);
    orc_expm_sym(n, T, +o->dtau, o->eT2p);
}
EXPORT void orc_set_hopping_exp(oracle_t *o, const double *eTmh, const double *eTph,
                                const double *eT2m, const double *eT2p) {
    memcpy(o->eTmh, eTmh, sizeof(double) * NN);
    memcpy(o->eTph, eTph, sizeof(double) * NN);
    memcpy(o->eT2m, eT2m, sizeof(double) * NN);
    memcpy(o->eT2p, eT2p, sizeof(double) * NN);
}
EXPORT void orc_get_hopping_exp(oracle_t *o, double *eTmh, double *eTph, double *eT2m, double *eT2p) {
    memcpy(eTmh, o->eTmh, sizeof(double) * NN);
    memcpy(eTph, o->eTph, sizeof(double) * NN);
    memcpy(eT2m, o->eT2m, sizeof(double) * NN);
    memcpy(eT2p, o->eT2p, sizeof(double) * NN);
}
EXPORT void orc_set_field(oracle_t *o, const int8_t *conf) { memcpy(o->conf, conf, (size_t)o->n * o->M); }
EXPORT void orc_get_field(oracle_t *o, int8_t *conf) { memcpy(conf, o->conf, (size_t)o->n * o->M); }
EXPORT void orc_set_uniforms(oracle_t *o, const double *u, int64_t count) {
    o->uniforms = u; o->n_uniforms = count; o->u_pos = 0; o->stream_exhausted = 0;
}
EXPORT void orc_set_seed_at(oracle_t *o, uint64_t seed, uint64_t pos) {
    o->uniforms = NULL; o->n_uniforms = 0; o->seed = seed; o->u_pos = (int64_t)pos;
}
EXPORT void orc_set_seed(oracle_t *o, uint64_t seed) {
    o->uniforms = NULL; o->n_uniforms = 0; o->seed = seed; o->u_pos = 0;
}
EXPORT void orc_set_log(oracle_t *o, uint8_t *dec, double *p, int64_t cap) {
    o->dec_log = dec; o->p_log = p; o->log_cap = cap; o->log_len = 0;
}

/* interaction_matrix_exp!, fields.jl:380-386 (density) and 429-438 (magnetic):
 * eV_b[i] = exp(s * power * alpha * x_i), s = +1 for block 1, -1 for block 2. */
static void interaction_exp(oracle_t *o, int slice, double power) {
    int n = o->n;
    const int8_t *x = o->conf + (size_t)(slice - 1) * n;
    for (int b = 0; b < o->flv; b++) {
        double s = (b == 0 ? 1.0 : -1.0) * power;
        for (int i = 0; i < n; i++) o->eV[b * n + i] = (s * x[i] > 0) ? o->ev_tab[1] : o->ev_tab[0];
    }
}

/* slice multiplies, stack.jl:319-367.  Hermitian hopping exponentials are
 * used through their (symmetric) data exactly as fallbacks.jl:67-68. */
static void mult_slice_left(oracle_t *o, int slice, double *Mx) {      /* M <- eT2 eV M */
    int n = o->n;
    interaction_exp(o, slice, 1.0);
    for (int b = 0; b < o->flv; b++) {
        mul_diag_left(n, BLK(o->tmp1, b), o->eV + b * n, BLK(Mx, b));
        gemm(n, 1, o->eT2m, 0, BLK(o->tmp1, b), BLK(Mx, b), o->work);
    }
}
static void mult_slice_right(oracle_t *o, int slice, double *Mx) {     /* M <- M eT2 eV */
    int n = o->n;
    interaction_exp(o, slice, 1.0);
    for (int b = 0; b < o->flv; b++) {
        gemm(n, 0, BLK(Mx, b), 1, o->eT2m, BLK(o->tmp1, b), o->work);
        mul_diag_right(n, BLK(Mx, b), BLK(o->tmp1, b), o->eV + b * n);
    }
}
static void mult_slice_inv_right(oracle_t *o, int slice, double *Mx) { /* M <- M eV^-1 eT2^-1 */
    int n = o->n;
    interaction_exp(o, slice, -1.0);
    for (int b = 0; b < o->flv; b++) {
        mul_diag_right(n, BLK(o->tmp1, b), BLK(Mx, b), o->eV + b * n);
        gemm(n, 0, BLK(o->tmp1, b), 1, o->eT2p, BLK(Mx, b), o->work);
    }
}
static void mult_slice_inv_left(oracle_t *o, int slice, double *Mx) {  /* M <- eV^-1 eT2^-1 M */
    int n = o->n;
    interaction_exp(o, slice, -1.0);
    for (int b = 0; b < o->flv; b++) {
        gemm(n, 1, o->eT2p, 0, BLK(Mx, b), BLK(o->tmp1, b), o->work);
        mul_diag_left(n, BLK(Mx, b), o->eV + b * n, BLK(o->tmp1, b));
    }
}
static void mult_daggered_slice_left(oracle_t *o, int slice, double *Mx) { /* M <- eV' eT2' M */
    int n = o->n;
    interaction_exp(o, slice, 1.0);
    for (int b = 0; b < o->flv; b++) {
        gemm(n, 1, o->eT2m, 0, BLK(Mx, b), BLK(o->tmp1, b), o->work);
        mul_diag_left(n, BLK(Mx, b), o->eV + b * n, BLK(o->tmp1, b));
    }
}

/* add_slice_sequence_left(mc, idx), stack.jl:377-393 */
static void add_slice_sequence_left(oracle_t *o, int idx) {
    int n = o->n;
    for (int b = 0; b < o->flv; b++) memcpy(BLK(o->curr_U, b), USTK(idx, b), sizeof(double) * NN);
    for (int s = o->rstart[idx - 1]; s <= o->rstop[idx - 1]; s++) mult_slice_left(o, s, o->curr_U);
    for (int b = 0; b < o->flv; b++) {
        mul_diag_right(n, BLK(o->tmp1, b), BLK(o->curr_U, b), DSTK(idx, b));
        udt_pivot(n, USTK(idx + 1, b), DSTK(idx + 1, b), BLK(o->tmp1, b), o->pivot, o->tempv, 1);
        gemm(n, 0, BLK(o->tmp1, b), 0, TSTK(idx, b), TSTK(idx + 1, b), o->work);
    }
}

/* add_slice_sequence_right(mc, idx), stack.jl:402-416 */
static void add_slice_sequence_right(oracle_t *o, int idx) {
    int n = o->n;
    for (int b = 0; b < o->flv; b++) memcpy(BLK(o->curr_U, b), USTK(idx + 1, b), sizeof(double) * NN);
    for (int s = o->rstop[idx - 1]; s >= o->rstart[idx - 1]; s--) mult_daggered_slice_left(o, s, o->curr_U);
    for (int b = 0; b < o->flv; b++) {
        mul_diag_right(n, BLK(o->tmp1, b), BLK(o->curr_U, b), DSTK(idx + 1, b));
        udt_pivot(n, USTK(idx, b), DSTK(idx, b), BLK(o->tmp1, b), o->pivot, o->tempv, 1);
        gemm(n, 0, BLK(o->tmp1, b), 0, TSTK(idx + 1, b), TSTK(idx, b), o->work);
    }
}

static void stack_slot_identity(oracle_t *o, int s) {
    for (int b = 0; b < o->flv; b++) {
        set_identity(o->n, USTK(s, b));
        set_identity(o->n, TSTK(s, b));
        for (int i = 0; i < o->n; i++) DSTK(s, b)[i] = 1.0;
    }
}

/* calculate_greens(mc, output), stack.jl:509-516 */
static void calculate_greens(oracle_t *o, double *out) {
    int n = o->n;
    for (int b = 0; b < o->flv; b++)
        calc_greens_udt(n, BLK(o->Ul, b), o->Dl + b * n, BLK(o->Tl, b), BLK(o->Ur, b), o->Dr + b * n,
                        BLK(o->Tr, b), BLK(out, b), o->pivot, o->tempv, o->work);
}

static void copy_slot_to(oracle_t *o, int s, double *U, double *D, double *T) {
    for (int b = 0; b < o->flv; b++) {
        memcpy(BLK(U, b), USTK(s, b), sizeof(double) * NN);
        memcpy(T + (size_t)b * NN, TSTK(s, b), sizeof(double) * NN);
        memcpy(D + (size_t)b * o->n, DSTK(s, b), sizeof(double) * o->n);
    }
}
static void set_udt_identity(oracle_t *o, double *U, double *D, double *T) {
    for (int b = 0; b < o->flv; b++) { set_identity(o->n, BLK(U, b)); set_identity(o->n, BLK(T, b)); }
    for (int i = 0; i < o->flv * o->n; i++) D[i] = 1.0;
}

/* build_stack, stack.jl:257-281 */
EXPORT void orc_build_stack_forward(oracle_t *o) {
    stack_slot_identity(o, 1);
    for (int i = 1; i <= o->K; i++) add_slice_sequence_left(o, i);
    o->current_slice = o->M + 1;
    o->current_range = o->K;
    o->direction = -1;
    copy_slot_to(o, o->K + 1, o->Ul, o->Dl, o->Tl);
    set_udt_identity(o, o->Ur, o->Dr, o->Tr);
    calculate_greens(o, o->greens);
}

/* reverse_build_stack, stack.jl:284-308 */
EXPORT void orc_reverse_build_stack(oracle_t *o) {
    stack_slot_identity(o, o->K + 1);
    for (int i = o->K; i >= 1; i--) add_slice_sequence_right(o, i);
    o->current_slice = 0;
    o->current_range = 1;
    o->direction = 1;
    set_udt_identity(o, o->Ul, o->Dl, o->Tl);
    copy_slot_to(o, 1, o->Ur, o->Dr, o->Tr);
    calculate_greens(o, o->greens);
}

/* wrap_greens!(mc, gf, curr_slice, direction), stack.jl:594-603 */
static void wrap_greens(oracle_t *o, double *gf, int curr_slice, int direction) {
    if (direction == -1) {
        mult_slice_inv_left(o, curr_slice - 1, gf);
        mult_slice_right(o, curr_slice - 1, gf);
    } else {
        mult_slice_left(o, curr_slice, gf);
        mult_slice_inv_right(o, curr_slice, gf);
    }
}

static void record_prop_error(oracle_t *o) {
    double diff = 0.0;
    size_t tot = (size_t)o->flv * NN;
    for (size_t i = 0; i < tot; i++) {
        double d = fabs(o->greens_temp[i] - o->greens[i]);
        if (d > diff || d != d) diff = d;
    }
    o->prop_err_last = diff;
    o->stabilizations++;
    if (diff > 1e-7) {
        if (o->prop_err_count < o->prop_err_cap) o->prop_err_log[o->prop_err_count] = diff;
        o->prop_err_count++;
        if (diff > o->prop_err_max) o->prop_err_max = diff;
    }
}

/* propagate(mc), stack.jl:605-730 */
EXPORT void orc_propagate(oracle_t *o) {
    size_t tot = (size_t)o->flv * NN;
    o->current_slice += o->direction;
    if (o->direction == 1) {
        if (o->current_slice == 1) {
            stack_slot_identity(o, 1);
        } else if (o->current_slice - 1 == o->rstop[o->current_range - 1]) {
            int idx = o->current_range;
            copy_slot_to(o, idx + 1, o->Ur, o->Dr, o->Tr);
            add_slice_sequence_left(o, idx);
            copy_slot_to(o, idx + 1, o->Ul, o->Dl, o->Tl);
            if (o->check_propagation_error) memcpy(o->greens_temp, o->greens, sizeof(double) * tot);
            wrap_greens(o, o->greens_temp, o->current_slice - 1, 1);
            calculate_greens(o, o->greens);
            if (o->check_propagation_error) record_prop_error(o);
            if (o->current_range == o->K) {
                o->direction = -1;
                orc_propagate(o);
            } else {
                o->current_range += 1;
            }
        } else {
            wrap_greens(o, o->greens, o->current_slice - 1, 1);
        }
    } else {
        if (o->current_slice == o->M) {
            stack_slot_identity(o, o->K + 1);
            wrap_greens(o, o->greens, o->current_slice + 1, -1);
        } else if (o->current_slice + 1 == o->rstart[o->current_range - 1]) {
            int idx = o->current_range;
            copy_slot_to(o, idx, o->Ul, o->Dl, o->Tl);
            add_slice_sequence_right(o, idx);
            copy_slot_to(o, idx, o->Ur, o->Dr, o->Tr);
            if (o->check_propagation_error) memcpy(o->greens_temp, o->greens, sizeof(double) * tot);
            calculate_greens(o, o->greens);
            if (o->check_propagation_error) record_prop_error(o);
            if (o->current_range == 1) {
                o->direction = 1;
                orc_propagate(o);
            } else {
                wrap_greens(o, o->greens, o->current_slice + 1, -1);
                o->current_range -= 1;
            }
        } else {
            wrap_greens(o, o->greens, o->current_slice + 1, -1);
        }
    }
}

/* Scalar decision arithmetic kept free of FMA contraction so that it rounds
 * as the reference's plain Julia expressions do (fields.jl:64,71-74,
 * local_updates.jl:31). */
__attribute__((optimize("fp-contract=off")))
static double detratio_p(oracle_t *o, int i, int slice, double *R, double *Dl) {
    int n = o->n;
    int8_t x = o->conf[(size_t)(slice - 1) * n + (i)];
    int ix = x > 0;
    if (o->flv == 1) {
        double d = o->d_first[ix];
        double g = o->greens[(size_t)i + (size_t)i * n];
        double r = 1.0 + d * (1.0 - g);
        R[0] = r; Dl[0] = d;
        double detratio = r * r;
        return o->bweight[ix] * detratio;
    } else {
        double d1 = o->d_first[ix], d2 = o->d_second[ix];
        double g1 = o->greens[(size_t)i + (size_t)i * n];
        double g2 = o->greens[NN + (size_t)i + (size_t)i * n];
        double r1 = 1.0 + d1 * (1.0 - g1);
        double r2 = 1.0 + d2 * (1.0 - g2);
        R[0] = r1; R[1] = r2; Dl[0] = d1; Dl[1] = d2;
        double detratio = r1 * r2;
        return 1.0 * detratio; /* exp(-0.0) * detratio */
    }
}

static double next_uniform(oracle_t *o) {
    if (o->uniforms) {
        if (o->u_pos >= o->n_uniforms) { o->stream_exhausted = 1; o->u_pos++; return 1.0; }
        return o->uniforms[o->u_pos++];
    }
    return orc_uniform_at(o->seed, (uint64_t)o->u_pos++);
}

/* update_greens!(cache, G, i, N), fields.jl:271-286 with vldiv22!, vsub!,
 * vmul!, vsubkron! (updates.jl:7-131) */
static void update_greens(oracle_t *o, int i, const double *R, const double *Dl) {
    int n = o->n;
    for (int b = 0; b < o->flv; b++) {
        double *G = BLK(o->greens, b);
        double invRD = Dl[b] / R[b];
        double *IG = o->IG + b * n, *g = o->Gr + b * n;
        for (int j = 0; j < n; j++) IG[j] = -G[j + (size_t)i * n];
        IG[i] += 1.0;
        for (int j = 0; j < n; j++) g[j] = invRD * G[i + (size_t)j * n];
        for (int l = 0; l < n; l++) {
            double gl = g[l];
            double *col = G + (size_t)l * n;
            for (int k = 0; k < n; k++) col[k] -= IG[k] * gl;
        }
    }
}

/* sweep_spatial(mc, m), local_updates.jl:23-60 */
EXPORT int64_t orc_sweep_spatial(oracle_t *o) {
    int slice = o->current_slice;
    int64_t acc = 0;
    double R[2], Dl[2];
    for (int i = 0; i < o->n; i++) {
        double p = detratio_p(o, i, slice, R, Dl);
        if (p < 0.0) o->negative_prob++;
        int accept;
        if (p > 1.0) {
            accept = 1;
            if (p - 1.0 < 1e-12) o->ties++;
        } else {
            double u = next_uniform(o);
            accept = u < p;
            if (fabs(u - p) < 1e-12) o->ties++;
        }
        o->proposed++;
        if (o->dec_log && o->log_len < o->log_cap) {
            o->dec_log[o->log_len] = (uint8_t)accept;
            if (o->p_log) o->p_log[o->log_len] = p;
            o->log_len++;
        }
        if (accept) {
            update_greens(o, i, R, Dl);
            o->conf[(size_t)(slice - 1) * o->n + i] *= -1;
            acc++;
        }
    }
    o->accepted += acc;
    return acc;
}

/* local_sweep(mc), local_updates.jl:7-14 */
EXPORT int64_t orc_sweep(oracle_t *o, int nsweeps) {
    int64_t acc = 0;
    for (int s = 0; s < nsweeps; s++)
        for (int k = 0; k < 2 * o->M; k++) {
            acc += orc_sweep_spatial(o);
            orc_propagate(o);
        }
    return acc;
}

/* calculate_greens(mc, slice, output), stack.jl:525-583 */
EXPORT void orc_calculate_greens_slice(oracle_t *o, int slice, double *out) {
    int n = o->n, sm = o->safe_mult, M = o->M;
    for (int b = 0; b < o->flv; b++) set_identity(n, BLK(o->curr_U, b));
    set_udt_identity(o, o->Ur, o->Dr, o->Tr);
    if (slice + 1 <= M) {
        for (int k = M; k >= slice + 1; k--) {
            mult_daggered_slice_left(o, k, o->curr_U);
            if (k % sm == 0) {
                for (int b = 0; b < o->flv; b++) {
                    mul_diag_right(n, BLK(o->tmp1, b), BLK(o->curr_U, b), o->Dr + b * n);
                    udt_pivot(n, BLK(o->curr_U, b), o->Dr + b * n, BLK(o->tmp1, b), o->pivot, o->tempv, 1);
                    memcpy(BLK(o->tmp2, b), BLK(o->Tr, b), sizeof(double) * NN);
                    gemm(n, 0, BLK(o->tmp1, b), 0, BLK(o->tmp2, b), BLK(o->Tr, b), o->work);
                }
            }
        }
        for (int b = 0; b < o->flv; b++) {
            mul_diag_right(n, BLK(o->tmp1, b), BLK(o->curr_U, b), o->Dr + b * n);
            udt_pivot(n, BLK(o->Ur, b), o->Dr + b * n, BLK(o->tmp1, b), o->pivot, o->tempv, 1);
            memcpy(BLK(o->tmp2, b), BLK(o->Tr, b), sizeof(double) * NN);
            gemm(n, 0, BLK(o->tmp1, b), 0, BLK(o->tmp2, b), BLK(o->Tr, b), o->work);
        }
    }
    for (int b = 0; b < o->flv; b++) set_identity(n, BLK(o->curr_U, b));
    set_udt_identity(o, o->Ul, o->Dl, o->Tl);
    if (slice >= 1) {
        for (int k = 1; k <= slice; k++) {
            mult_slice_left(o, k, o->curr_U);
            if (k % sm == 0) {
                for (int b = 0; b < o->flv; b++) {
                    mul_diag_right(n, BLK(o->tmp1, b), BLK(o->curr_U, b), o->Dl + b * n);
                    udt_pivot(n, BLK(o->curr_U, b), o->Dl + b * n, BLK(o->tmp1, b), o->pivot, o->tempv, 1);
                    memcpy(BLK(o->tmp2, b), BLK(o->Tl, b), sizeof(double) * NN);
                    gemm(n, 0, BLK(o->tmp1, b), 0, BLK(o->tmp2, b), BLK(o->Tl, b), o->work);
                }
            }
        }
        for (int b = 0; b < o->flv; b++) {
            mul_diag_right(n, BLK(o->tmp1, b), BLK(o->curr_U, b), o->Dl + b * n);
            udt_pivot(n, BLK(o->Ul, b), o->Dl + b * n, BLK(o->tmp1, b), o->pivot, o->tempv, 1);
            memcpy(BLK(o->tmp2, b), BLK(o->Tl, b), sizeof(double) * NN);
            gemm(n, 0, BLK(o->tmp1, b), 0, BLK(o->tmp2, b), BLK(o->Tl, b), o->work);
        }
    }
    calculate_greens(o, out);
}

/* greens(mc): e^{+dtau T/2} G e^{-dtau T/2}, greens.jl:114-125 */
/* global_update(mc, model, field) (global_updates.jl:191-207) for the proposal `proposal`
 * (the new field.conf; the old one plays temp_conf):
 *   propose_global_from_conf (:136-169): tempvf = Dl, the inverse D of the last
 *     calculate_greens; inv_det(mc, current_slice - 1 = 0) (:64-121): the chain from scratch
 *     at slice 0 with the proposal, whose calculate_inv_greens_udt (:25-47) leaves the new D
 *     in Dr (the same steps as calculate_greens up to its last UDT); detratio = prod over
 *     all entries of tempvf[i] * Dr[i], squared for one unique flavour; dE = E_b(new) - E_b(old)
 *     with energy_boson = alpha sum(conf) (density) or 0 (magnetic) (fields.jl:395, 451);
 *   p = |exp(-dE) detratio|; accepted iff p > 1 || rand() < p (a uniform only when p <= 1);
 *   accept_global! (:171-188): reverse_build_stack + propagate; otherwise the old conf is
 *   restored.
 * Requires current_slice == 1 and direction == 1 (the state after a sweep or a build).
 * Returns 1 if accepted, 0 if rejected, -1 if the state is wrong; *p_out = p. */
EXPORT int orc_global_update(oracle_t *o, const int8_t *proposal, double *p_out) {
    int n = o->n, M = o->M, flv = o->flv;
    if (o->current_slice != 1 || o->direction != 1) return -1;
    double *tempvf = malloc(sizeof(double) * flv * n);
    int8_t *old = malloc((size_t)n * M);
    memcpy(tempvf, o->Dl, sizeof(double) * flv * n);
    memcpy(old, o->conf, (size_t)n * M);
    memcpy(o->conf, proposal, (size_t)n * M);
    orc_calculate_greens_slice(o, 0, o->greens_temp);
    double detratio = 1.0;
    for (int i = 0; i < flv * n; i++) detratio *= tempvf[i] * o->Dr[i];
    if (flv == 1) detratio = detratio * detratio;
    double dE = 0.0;
    if (flv == 1) {
        int64_t s_new = 0, s_old = 0;
        for (size_t i = 0; i < (size_t)n * M; i++) { s_new += o->conf[i]; s_old += old[i]; }
        dE = o->alpha * (double)s_new - o->alpha * (double)s_old;
    }
    double p = fabs(exp(-dE) * detratio);
    *p_out = p;
    int accept = p > 1.0 || next_uniform(o) < p;
    if (accept) {
        orc_reverse_build_stack(o);
        orc_propagate(o);
    } else {
        memcpy(o->conf, old, (size_t)n * M);
    }
    free(tempvf);
    free(old);
    return accept;
}

EXPORT void orc_get_greens(oracle_t *o, double *out, int physical) {
    int n = o->n;
    if (!physical) { memcpy(out, o->greens, sizeof(double) * o->flv * NN); return; }
    for (int b = 0; b < o->flv; b++) {
        gemm(n, 0, BLK(o->greens, b), 1, o->eTmh, BLK(o->tmp2, b), o->work);
        gemm(n, 1, o->eTph, 0, BLK(o->tmp2, b), BLK(out, b), o->work);
    }
}
EXPORT void orc_set_greens(oracle_t *o, const double *G) { memcpy(o->greens, G, sizeof(double) * o->flv * NN); }

EXPORT void orc_get_state(oracle_t *o, int *cs, int *dir, int *cr) {
    *cs = o->current_slice; *dir = o->direction; *cr = o->current_range;
}

/* statistics: [accepted, proposed, uniforms consumed, negative, ties,
 *              prop_err_count, stabilizations, stream_exhausted, log_len] */
EXPORT void orc_get_counters(oracle_t *o, int64_t *c) {
    c[0] = o->accepted; c[1] = o->proposed; c[2] = o->u_pos; c[3] = o->negative_prob;
    c[4] = o->ties; c[5] = o->prop_err_count; c[6] = o->stabilizations; c[7] = o->stream_exhausted;
    c[8] = o->log_len;
}
EXPORT double orc_prop_err_max(oracle_t *o) { return o->prop_err_max; }
EXPORT double orc_prop_err_last(oracle_t *o) { return o->prop_err_last; }

/* copies of stack slots (1-based s) for the build == reverse-build test */
EXPORT void orc_get_stack_slot(oracle_t *o, int s, double *U, double *D, double *T) {
    copy_slot_to(o, s, U, D, T);
}

/* --- low-level entry points for the linear-algebra unit tests ------------- */
EXPORT void orc_udt(int n, double *input, double *U, double *D, int *pivot, int apply_pivot) {
    double *temp = malloc(sizeof(double) * n);
    udt_pivot(n, U, D, input, pivot, temp, apply_pivot);
    free(temp);
}
EXPORT void orc_rdivp(int n, double *A, const double *T, const int *pivot) {
    double *O = malloc(sizeof(double) * (size_t)n * n);
    rdivp(n, A, T, O, pivot);
    free(O);
}
EXPORT void orc_calc_greens_udt(int n, double *Ul, double *Dl, double *Tl, double *Ur, double *Dr,
                                double *Tr, double *G) {
    int *pivot = malloc(sizeof(int) * n);
    double *temp = malloc(sizeof(double) * n);
    double *work = malloc(sizeof(double) * 2 * (size_t)n * n);
    calc_greens_udt(n, Ul, Dl, Tl, Ur, Dr, Tr, G, pivot, temp, work);
    free(pivot); free(temp); free(work);
}
EXPORT void orc_gemm(int n, int ta, const double *A, int tb, const double *B, double *C) {
    double *work = malloc(sizeof(double) * 2 * (size_t)n * n);
    gemm(n, ta, A, tb, B, C, work);
    free(work);
}
