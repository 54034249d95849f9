/*
 * CPU restatement of the reference's forward H2 product in C with OpenMP -- TEST INFRASTRUCTURE
 * ONLY: the CPU baseline that bench.py times on the host cores beside the GPU, and a second
 * checker of the product at small sizes.  Nothing in the product loads it.
 *
 * It runs the reference traversal of mul!(y, A, x) (src/maps/abstracth2map.jl:42-86) in FP64 on
 * the same core data the GPU path receives, for real Laplace maps with Toeplitz couplings:
 *   near   y += N x, the stored block-sparse operator (src/nearinteractions/nearmul.jl:8-14);
 *   P2M    m_leaf = C_leaf x_leaf (src/moments/create/linkmomenttocollection.jl:56-63);
 *   M2M    m_parent = sum over children T_sector m_child, first child overwrites
 *          (src/farinteractions/traversalhelpers.jl:49-75, src/maps/leveltranslator.jl:131-158);
 *   M2L    per receiving box: crop(ifft(sum_partners D_op .* fft(pad(m_partner))))
 *          (src/moments/toeplitz/toeplitzmul.jl:86-305: modulate_in, Hadamard, modulate_out);
 *   L2L    loc_child = (I2O wrote it ? loc_child : 0) + T_sector^T loc_parent
 *          (src/farinteractions/disaggregation.jl:244-270);
 *   L2P    y_leaf += C_leaf^T loc_leaf for initialised leaves
 *          (src/farinteractions/traversalhelpers.jl:103-112).
 * The reference parallelises each level with tasks over node chunks (OhMyThreads @tasks); here
 * each level is one OpenMP loop over its boxes.  The 3-D DFTs of the odd length L = 2p - 1 are
 * separable direct DFTs over the nonzero (padded) or kept (cropped) entries.
 * Near values and leaf collections are computed at set-up (not timed), from the points, with the
 * Laplace kernel 1/(4 pi r) (0 at r = 0) and tensor Lagrange bases on the box nodes.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int64_t n;                  /* points (one tree for test and trial)                      */
    const double *points;       /* [n][3] original order                                      */
    const int64_t *perm;        /* [n] tree position -> original index                        */
    int32_t n_levels, n_boxes;
    const int32_t *level_offsets, *parent, *sector;
    const int64_t *begin, *end; /* tree-order dof range of each box                           */
    const uint8_t *is_leaf;
    const double *center;       /* [n_boxes][3]                                               */
    const double *halfwidth;    /* [n_boxes]                                                  */
    int32_t p;
    const double *nodes;        /* [p] reference nodes in [-1, 1]                             */
    const double *transfers;    /* [n_levels][8] K x K column-major, child (sector) -> parent  */
    int64_t n_pairs;
    const int32_t *pair_test, *pair_trial, *pair_op;
    int64_t n_ops;
    const double *spectra;      /* [n_ops][L^3] complex (re, im), x fastest: D = fft(kappa)   */
    int64_t n_near;
    const int32_t *near_test, *near_trial;
    const int64_t *near_offset; /* block b: column-major nr x nc at near_offset[b]             */
    int64_t n_near_values;
} h2cpu_desc;

typedef struct {
    h2cpu_desc d;
    int K, L, L3;
    double *coll;      /* [n][K]: column j of the K x n collection                         */
    double *near;      /* the near values                                                  */
    int64_t *near_ptr; /* CSR of the near blocks by test leaf (box index)                  */
    int32_t *near_blk;
    int64_t *m2l_ptr;  /* CSR of the pairs by receiving (test) box                         */
    int32_t *m2l_src, *m2l_op;
    int32_t *child_begin, *child_end;
    uint8_t *has_m2l;
    double *tw_re, *tw_im; /* exp(-2 pi i k / L) */
    double *xs, *ys, *mom, *loc;
    double *spec;      /* [n_boxes][L^3] complex spectra of the source boxes                */
} h2cpu;

#define FOUR_PI 12.566370614359172

static double lagrange(const double *nodes, int p, int k, double t) {
    double v = 1.0;
    for (int m = 0; m < p; ++m)
        if (m != k) v *= (t - nodes[m]) / (nodes[k] - nodes[m]);
    return v;
}

static int64_t *csr(int nb, int64_t n, const int32_t *key, int32_t **order_out) {
    int64_t *ptr = calloc((size_t)nb + 1, sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) ptr[key[i] + 1]++;
    for (int b = 0; b < nb; ++b) ptr[b + 1] += ptr[b];
    int64_t *fill = malloc((size_t)nb * sizeof(int64_t));
    memcpy(fill, ptr, (size_t)nb * sizeof(int64_t));
    int32_t *order = malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    for (int64_t i = 0; i < n; ++i) order[fill[key[i]]++] = (int32_t)i;
    free(fill);
    *order_out = order;
    return ptr;
}

void *h2cpu_setup(const h2cpu_desc *desc) {
    if (desc->p < 2 || desc->p > 6) return NULL; /* L = 2p - 1 <= 11 */
    h2cpu *h = calloc(1, sizeof(h2cpu));
    h->d = *desc;
    const h2cpu_desc *d = &h->d;
    const int p = d->p, K = p * p * p, L = 2 * p - 1;
    h->K = K;
    h->L = L;
    h->L3 = L * L * L;
    /* leaf collections */
    h->coll = malloc((size_t)d->n * K * sizeof(double));
#pragma omp parallel for schedule(dynamic, 16)
    for (int b = 0; b < d->n_boxes; ++b) {
        if (!d->is_leaf[b]) continue;
        for (int64_t j = d->begin[b]; j < d->end[b]; ++j) {
            const double *x = d->points + 3 * d->perm[j];
            double lx[8], ly[8], lz[8];
            for (int k = 0; k < p; ++k) {
                lx[k] = lagrange(d->nodes, p, k, (x[0] - d->center[3 * b]) / d->halfwidth[b]);
                ly[k] = lagrange(d->nodes, p, k, (x[1] - d->center[3 * b + 1]) / d->halfwidth[b]);
                lz[k] = lagrange(d->nodes, p, k, (x[2] - d->center[3 * b + 2]) / d->halfwidth[b]);
            }
            double *c = h->coll + j * K;
            for (int kz = 0; kz < p; ++kz)
                for (int ky = 0; ky < p; ++ky)
                    for (int kx = 0; kx < p; ++kx) c[kx + p * (ky + p * kz)] = lx[kx] * ly[ky] * lz[kz];
        }
    }
    /* near blocks */
    h->near = malloc((size_t)(d->n_near_values > 0 ? d->n_near_values : 1) * sizeof(double));
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < d->n_near; ++q) {
        const int bt = d->near_test[q], bs = d->near_trial[q];
        const int64_t m = d->end[bt] - d->begin[bt], nc = d->end[bs] - d->begin[bs];
        double *B = h->near + d->near_offset[q];
        for (int64_t j = 0; j < nc; ++j) {
            const double *y = d->points + 3 * d->perm[d->begin[bs] + j];
            for (int64_t i = 0; i < m; ++i) {
                const double *x = d->points + 3 * d->perm[d->begin[bt] + i];
                const double dx = x[0] - y[0], dy = x[1] - y[1], dz = x[2] - y[2];
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                B[i + m * j] = r > 0.0 ? 1.0 / (FOUR_PI * r) : 0.0;
            }
        }
    }
    h->near_ptr = csr(d->n_boxes, d->n_near, d->near_test, &h->near_blk);
    h->m2l_ptr = csr(d->n_boxes, d->n_pairs, d->pair_test, &h->m2l_src);
    /* m2l_src holds pair indices; resolve to (partner box, op) */
    h->m2l_op = malloc((size_t)(d->n_pairs > 0 ? d->n_pairs : 1) * sizeof(int32_t));
    for (int64_t i = 0; i < d->n_pairs; ++i) {
        const int32_t k = h->m2l_src[i];
        h->m2l_op[i] = d->pair_op[k];
        h->m2l_src[i] = d->pair_trial[k];
    }
    h->has_m2l = calloc((size_t)d->n_boxes, 1);
    for (int64_t k = 0; k < d->n_pairs; ++k) h->has_m2l[d->pair_test[k]] = 1;
    h->child_begin = malloc((size_t)d->n_boxes * sizeof(int32_t));
    h->child_end = malloc((size_t)d->n_boxes * sizeof(int32_t));
    for (int b = 0; b < d->n_boxes; ++b) h->child_begin[b] = h->child_end[b] = 0;
    for (int b = d->n_boxes - 1; b >= 1; --b) {
        const int q = d->parent[b];
        if (q < 0) continue;
        if (h->child_end[q] == 0) h->child_end[q] = b + 1;
        h->child_begin[q] = b;
    }
    h->tw_re = malloc((size_t)L * sizeof(double));
    h->tw_im = malloc((size_t)L * sizeof(double));
    for (int k = 0; k < L; ++k) {
        h->tw_re[k] = cos(-2.0 * M_PI * k / L);
        h->tw_im[k] = sin(-2.0 * M_PI * k / L);
    }
    h->xs = malloc((size_t)d->n * sizeof(double));
    h->ys = malloc((size_t)d->n * sizeof(double));
    h->mom = malloc((size_t)d->n_boxes * K * sizeof(double));
    h->loc = malloc((size_t)d->n_boxes * K * sizeof(double));
    h->spec = malloc((size_t)d->n_boxes * h->L3 * 2 * sizeof(double));
    return h;
}

/* S = fft3(pad(m)), m p^3 real at the corner, S L^3 complex (x fastest) */
static void fft_in(const h2cpu *h, const double *m, double *S) {
    const int p = h->d.p, L = h->L;
    double A[2 * 11 * 11 * 11], B[2 * 11 * 11 * 11]; /* L <= 11 */
    for (int z = 0; z < p; ++z)
        for (int y = 0; y < p; ++y)
            for (int f = 0; f < L; ++f) {
                double re = 0, im = 0;
                for (int x = 0; x < p; ++x) {
                    const int w = (f * x) % L;
                    re += m[x + p * (y + p * z)] * h->tw_re[w];
                    im += m[x + p * (y + p * z)] * h->tw_im[w];
                }
                A[2 * (f + L * (y + p * z))] = re;
                A[2 * (f + L * (y + p * z)) + 1] = im;
            }
    for (int z = 0; z < p; ++z)
        for (int g = 0; g < L; ++g)
            for (int f = 0; f < L; ++f) {
                double re = 0, im = 0;
                for (int y = 0; y < p; ++y) {
                    const int w = (g * y) % L;
                    const double ar = A[2 * (f + L * (y + p * z))], ai = A[2 * (f + L * (y + p * z)) + 1];
                    re += ar * h->tw_re[w] - ai * h->tw_im[w];
                    im += ar * h->tw_im[w] + ai * h->tw_re[w];
                }
                B[2 * (f + L * (g + L * z))] = re;
                B[2 * (f + L * (g + L * z)) + 1] = im;
            }
    for (int e = 0; e < L; ++e)
        for (int g = 0; g < L; ++g)
            for (int f = 0; f < L; ++f) {
                double re = 0, im = 0;
                for (int z = 0; z < p; ++z) {
                    const int w = (e * z) % L;
                    const double br = B[2 * (f + L * (g + L * z))], bi = B[2 * (f + L * (g + L * z)) + 1];
                    re += br * h->tw_re[w] - bi * h->tw_im[w];
                    im += br * h->tw_im[w] + bi * h->tw_re[w];
                }
                S[2 * (f + L * (g + L * e))] = re;
                S[2 * (f + L * (g + L * e)) + 1] = im;
            }
}

/* out = crop(ifft3(O)) (normalised), real part */
static void fft_out(const h2cpu *h, const double *O, double *out) {
    const int p = h->d.p, L = h->L;
    double A[2 * 11 * 11 * 11], B[2 * 11 * 11 * 11];
    for (int z = 0; z < p; ++z)
        for (int g = 0; g < L; ++g)
            for (int f = 0; f < L; ++f) {
                double re = 0, im = 0;
                for (int e = 0; e < L; ++e) {
                    const int w = (L - (e * z) % L) % L; /* exp(+2 pi i e z / L) */
                    const double orr = O[2 * (f + L * (g + L * e))], oi = O[2 * (f + L * (g + L * e)) + 1];
                    re += orr * h->tw_re[w] - oi * h->tw_im[w];
                    im += orr * h->tw_im[w] + oi * h->tw_re[w];
                }
                A[2 * (f + L * (g + L * z))] = re;
                A[2 * (f + L * (g + L * z)) + 1] = im;
            }
    for (int z = 0; z < p; ++z)
        for (int y = 0; y < p; ++y)
            for (int f = 0; f < L; ++f) {
                double re = 0, im = 0;
                for (int g = 0; g < L; ++g) {
                    const int w = (L - (g * y) % L) % L;
                    const double ar = A[2 * (f + L * (g + L * z))], ai = A[2 * (f + L * (g + L * z)) + 1];
                    re += ar * h->tw_re[w] - ai * h->tw_im[w];
                    im += ar * h->tw_im[w] + ai * h->tw_re[w];
                }
                B[2 * (f + L * (y + p * z))] = re;
                B[2 * (f + L * (y + p * z)) + 1] = im;
            }
    const double sc = 1.0 / (double)h->L3;
    for (int z = 0; z < p; ++z)
        for (int y = 0; y < p; ++y)
            for (int x = 0; x < p; ++x) {
                double re = 0;
                for (int f = 0; f < L; ++f) {
                    const int w = (L - (f * x) % L) % L;
                    re += B[2 * (f + L * (y + p * z))] * h->tw_re[w] - B[2 * (f + L * (y + p * z)) + 1] * h->tw_im[w];
                }
                out[x + p * (y + p * z)] = sc * re;
            }
}

/* y = A x (forward, one right-hand side, original order) */
void h2cpu_mul(void *handle, const double *x, double *y) {
    h2cpu *h = handle;
    const h2cpu_desc *d = &h->d;
    const int K = h->K, L3 = h->L3;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < d->n; ++i) {
        h->xs[i] = x[d->perm[i]];
        h->ys[i] = 0.0;
    }
    /* near: one test leaf per iteration (rows disjoint) */
#pragma omp parallel for schedule(dynamic, 32)
    for (int b = 0; b < d->n_boxes; ++b) {
        for (int64_t q = h->near_ptr[b]; q < h->near_ptr[b + 1]; ++q) {
            const int blk = h->near_blk[q];
            const int bs = d->near_trial[blk];
            const int64_t m = d->end[b] - d->begin[b], nc = d->end[bs] - d->begin[bs];
            const double *B = h->near + d->near_offset[blk];
            double *yo = h->ys + d->begin[b];
            const double *xi = h->xs + d->begin[bs];
            for (int64_t j = 0; j < nc; ++j) {
                const double xv = xi[j];
                const double *col = B + m * j;
                for (int64_t i = 0; i < m; ++i) yo[i] += col[i] * xv;
            }
        }
    }
    /* upward pass */
    for (int l = d->n_levels - 1; l >= 0; --l) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int b = d->level_offsets[l]; b < d->level_offsets[l + 1]; ++b) {
            double *m = h->mom + (int64_t)b * K;
            if (d->is_leaf[b]) {
                for (int k = 0; k < K; ++k) m[k] = 0.0;
                for (int64_t j = d->begin[b]; j < d->end[b]; ++j) {
                    const double xv = h->xs[j];
                    const double *c = h->coll + j * K;
                    for (int k = 0; k < K; ++k) m[k] += c[k] * xv;
                }
            } else {
                for (int k = 0; k < K; ++k) m[k] = 0.0;
                for (int c = h->child_begin[b]; c < h->child_end[b]; ++c) {
                    const double *T = d->transfers + ((int64_t)(l + 1) * 8 + d->sector[c]) * K * K;
                    const double *mc = h->mom + (int64_t)c * K;
                    for (int j = 0; j < K; ++j) {
                        const double v = mc[j];
                        const double *col = T + (int64_t)K * j;
                        for (int k = 0; k < K; ++k) m[k] += col[k] * v;
                    }
                }
            }
        }
    }
    /* couplings, level by level: spectra of the level's boxes, then the receiving boxes */
    for (int l = 0; l < d->n_levels; ++l) {
        const int b0 = d->level_offsets[l], b1 = d->level_offsets[l + 1];
        int any = 0;
        for (int b = b0; b < b1 && !any; ++b) any = h->m2l_ptr[b + 1] > h->m2l_ptr[b];
        if (!any) continue;
#pragma omp parallel for schedule(dynamic, 16)
        for (int b = b0; b < b1; ++b) fft_in(h, h->mom + (int64_t)b * K, h->spec + (int64_t)(b - b0) * 2 * L3);
#pragma omp parallel for schedule(dynamic, 16)
        for (int t = b0; t < b1; ++t) {
            if (h->m2l_ptr[t + 1] == h->m2l_ptr[t]) continue;
            double O[2 * 11 * 11 * 11];
            for (int f = 0; f < 2 * L3; ++f) O[f] = 0.0;
            for (int64_t q = h->m2l_ptr[t]; q < h->m2l_ptr[t + 1]; ++q) {
                const double *D = d->spectra + (int64_t)h->m2l_op[q] * 2 * L3;
                const double *S = h->spec + (int64_t)(h->m2l_src[q] - b0) * 2 * L3;
                for (int f = 0; f < L3; ++f) {
                    O[2 * f] += D[2 * f] * S[2 * f] - D[2 * f + 1] * S[2 * f + 1];
                    O[2 * f + 1] += D[2 * f] * S[2 * f + 1] + D[2 * f + 1] * S[2 * f];
                }
            }
            fft_out(h, O, h->loc + (int64_t)t * K);
        }
    }
    /* downward pass: initialised boxes get L2L from their parent; leaves L2P */
    uint8_t *init = calloc((size_t)d->n_boxes, 1);
    for (int l = 0; l < d->n_levels; ++l) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int b = d->level_offsets[l]; b < d->level_offsets[l + 1]; ++b) {
            double *lo = h->loc + (int64_t)b * K;
            const int q = d->parent[b];
            const int up = q >= 0 && init[q];
            if (up) {
                if (!h->has_m2l[b])
                    for (int k = 0; k < K; ++k) lo[k] = 0.0;
                const double *T = d->transfers + ((int64_t)l * 8 + d->sector[b]) * K * K;
                const double *lp = h->loc + (int64_t)q * K;
                for (int j = 0; j < K; ++j) {
                    const double *col = T + (int64_t)K * j;
                    double s = 0.0;
                    for (int k = 0; k < K; ++k) s += col[k] * lp[k];
                    lo[j] += s;
                }
            }
            init[b] = h->has_m2l[b] || up;
            if (init[b] && d->is_leaf[b])
                for (int64_t j = d->begin[b]; j < d->end[b]; ++j) {
                    const double *c = h->coll + j * K;
                    double s = 0.0;
                    for (int k = 0; k < K; ++k) s += c[k] * lo[k];
                    h->ys[j] += s;
                }
        }
    }
    free(init);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < d->n; ++i) y[d->perm[i]] = h->ys[i];
}

/* threads the OpenMP loops use */
int h2cpu_threads(void) { return omp_get_max_threads(); }

void h2cpu_free(void *handle) {
    h2cpu *h = handle;
    if (!h) return;
    free(h->coll), free(h->near), free(h->near_ptr), free(h->near_blk), free(h->m2l_ptr), free(h->m2l_src);
    free(h->m2l_op), free(h->child_begin), free(h->child_end), free(h->has_m2l), free(h->tw_re), free(h->tw_im);
    free(h->xs), free(h->ys), free(h->mom), free(h->loc), free(h->spec);
    free(h);
}
