/*
 * ORACLE — test infrastructure only; never linked into or called by the product.
 *
 * Plain-C restatement of the hop tables and the H·v loop nest of
 * NuclearToolkit.jl (paths relative to the reference repository root):
 *   TF_connectable / TF_connectable_1   src/ShellModel/shellmodel_main.jl:850-867
 *   calc_phase! / calc_phase_1!         shellmodel_main.jl:869-897
 *   bisearch! (descending)              shellmodel_main.jl:1068-1083
 *   prep_2bjumps                        shellmodel_main.jl:1115-1137
 *   prep_pn / calc_1b_jumps!            shellmodel_main.jl:1139-1199
 *   operate_H!                          src/ShellModel/lanczos_methods.jl:153-248
 * The tables keep the reference's structure: one T1info list per (block, SD)
 * with entries only for i <= f (coef halved on the diagonal), and one proton and
 * one neutron ifph hop list per (block pair bi<=bf, operator).  operate_H keeps
 * the owner-computes pn1 / pn2 / pp / nn / one-body segments in that order, so
 * its summation order per output element is the reference's.
 * Indices are 0-based here (the reference is 1-based).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef struct { int32_t i, f; int8_t ph; } ifph;        /* ph = 1 <=> phase == -1 */
typedef struct { int32_t f; double coef; } t1info;

typedef struct { ifph* a; int64_t n, cap; } ifvec;
typedef struct { t1info* a; int64_t n, cap; } t1vec;

typedef struct {
    int nb;
    int64_t *pbits, *poff, *nbits, *noff, *tdims, *Mps;
    int nop[2]; int64_t* ops[2]; double* V[2];             /* pp / nn operator lists */
    int ndm; int64_t* dms; int64_t* pnoff; int64_t* pnops; double* pnV;
    int norb_p, norb_n; double *spe_p, *spe_n, *jp, *jn;  /* jp: per proton SD (global) x norb */
    /* tables */
    t1vec* t1[2];                                          /* per global SD of the species */
    int* nbf; int** bfs; int* nbis; int** bis;             /* block connectivity */
    int** bf_dm;                                           /* ΔM index per (bi, j) */
    ifvec** plist; ifvec** nlist;                          /* [bi][j*nops_k + op] */
} orc_t;

static void ifpush(ifvec* v, ifph e) {
    if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 8; v->a = (ifph*)realloc(v->a, v->cap * sizeof(ifph)); }
    v->a[v->n++] = e;
}
static void t1push(t1vec* v, t1info e) {
    if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 8; v->a = (t1info*)realloc(v->a, v->cap * sizeof(t1info)); }
    v->a[v->n++] = e;
}

static int nthreads_req = 0;
void orc_set_threads(int n) { nthreads_req = n; if (n > 0) omp_set_num_threads(n); }

/* bisearch! on a strictly descending array; -1 if absent */
static int64_t bisearch_desc(const int64_t* v, int64_t n, int64_t t) {
    int64_t lo = 0, hi = n - 1;
    while (lo <= hi) {
        int64_t mid = (lo + hi) / 2;           /* div(lo+hi,2) on 1-based is the same split */
        if (v[mid] > t) lo = mid + 1;
        else if (v[mid] < t) hi = mid - 1;
        else return mid;
    }
    return -1;
}

static inline int popc(int64_t x) { return __builtin_popcountll((uint64_t)x); }

/* TF_connectable + calc_phase! : returns 0 if not connectable */
static int hop2(int64_t Phi, const int64_t* op, int64_t* fin, int* neg) {
    int64_t a = op[0], b = op[1], c = op[2], d = op[3];
    if (c != (c & Phi)) return 0;
    if (d != (d & Phi)) return 0;
    int64_t APhi = (~d) & ((~c) & Phi);
    if (a != (a & (~APhi))) return 0;
    if (b != (b & (~APhi))) return 0;
    int e = popc(Phi & ~(c - 1));
    int64_t P = (~c) & Phi;
    e += popc(P & ~(d - 1));
    P = (~d) & P;
    e += popc(P & ~(b - 1));
    P = b | P;
    e += popc(P & ~(a - 1));
    P = a | P;
    *fin = P; *neg = e & 1;
    return 1;
}

/* TF_connectable_1 + calc_phase_1! (note the initial +1 of the reference) */
static int hop1(int64_t Phi, int64_t a, int64_t c, int64_t* fin, int* neg) {
    if (c != (c & Phi)) return 0;
    if (a != (a & (~((~c) & Phi)))) return 0;
    int e = 1 + popc(Phi & ~(c - 1));
    int64_t P = (~c) & Phi;
    e += popc(P & ~(a - 1));
    *fin = a | P; *neg = e & 1;
    return 1;
}

static int64_t dm_index(const orc_t* h, int64_t dm) {
    for (int k = 0; k < h->ndm; k++) if (h->dms[k] == dm) return k;   /* dms descending, small */
    return -1;
}

static int64_t* dup64(const int64_t* s, int64_t n) { int64_t* d = (int64_t*)malloc((n > 0 ? n : 1) * 8); if (n) memcpy(d, s, n * 8); return d; }
static double* dupd(const double* s, int64_t n) { double* d = (double*)malloc((n > 0 ? n : 1) * 8); if (n) memcpy(d, s, n * 8); return d; }

void* orc_build(int nb, const int64_t* pbits, const int64_t* poff, const int64_t* nbits, const int64_t* noff,
                const int64_t* tdims, const int64_t* Mps,
                int npp, const int64_t* ppops, const double* ppV, int nnn, const int64_t* nnops, const double* nnV,
                int ndm, const int64_t* dms, const int64_t* pnoff, const int64_t* pnops, const double* pnV,
                const double* spe_p, const double* spe_n, const double* jp, const double* jn,
                int norb_p, int norb_n) {
    orc_t* h = (orc_t*)calloc(1, sizeof(orc_t));
    h->nb = nb;
    h->pbits = dup64(pbits, poff[nb]); h->poff = dup64(poff, nb + 1);
    h->nbits = dup64(nbits, noff[nb]); h->noff = dup64(noff, nb + 1);
    h->tdims = dup64(tdims, nb + 1); h->Mps = dup64(Mps, nb);
    h->nop[0] = npp; h->ops[0] = dup64(ppops, 4 * (int64_t)npp); h->V[0] = dupd(ppV, npp);
    h->nop[1] = nnn; h->ops[1] = dup64(nnops, 4 * (int64_t)nnn); h->V[1] = dupd(nnV, nnn);
    h->ndm = ndm; h->dms = dup64(dms, ndm); h->pnoff = dup64(pnoff, ndm + 1);
    h->pnops = dup64(pnops, 4 * pnoff[ndm]); h->pnV = dupd(pnV, pnoff[ndm]);
    h->norb_p = norb_p; h->norb_n = norb_n;
    h->spe_p = dupd(spe_p, norb_p); h->spe_n = dupd(spe_n, norb_n);
    h->jp = dupd(jp, poff[nb] * norb_p); h->jn = dupd(jn, noff[nb] * norb_n);

    /* prep_2bjumps for pp (species 0 over pbits) and nn (species 1 over nbits) */
    for (int s = 0; s < 2; s++) {
        const int64_t* bits = s == 0 ? h->pbits : h->nbits;
        const int64_t* off = s == 0 ? h->poff : h->noff;
        h->t1[s] = (t1vec*)calloc(off[nb] > 0 ? off[nb] : 1, sizeof(t1vec));
        #pragma omp parallel for schedule(dynamic, 1)
        for (int b = 0; b < nb; b++) {
            int64_t n = off[b + 1] - off[b];
            const int64_t* B = bits + off[b];
            for (int64_t i = 0; i < n; i++) {
                for (int k = 0; k < h->nop[s]; k++) {
                    if (h->V[s][k] == 0.0) continue;
                    int64_t fin; int neg;
                    if (!hop2(B[i], h->ops[s] + 4 * k, &fin, &neg)) continue;
                    int64_t r = bisearch_desc(B, n, fin);
                    if (r >= 0 && i <= r) {
                        double coef = h->V[s][k] * (neg ? -1.0 : 1.0) * (i == r ? 0.5 : 1.0);
                        t1info e = {(int32_t)r, coef};
                        t1push(&h->t1[s][off[b] + i], e);
                    }
                }
            }
        }
    }

    /* prep_pn: block connectivity */
    h->nbf = (int*)calloc(nb, sizeof(int)); h->nbis = (int*)calloc(nb, sizeof(int));
    h->bfs = (int**)calloc(nb, sizeof(int*)); h->bis = (int**)calloc(nb, sizeof(int*));
    h->bf_dm = (int**)calloc(nb, sizeof(int*));
    for (int i = 0; i < nb; i++) { h->bfs[i] = (int*)malloc(nb * sizeof(int)); h->bis[i] = (int*)malloc(nb * sizeof(int)); h->bf_dm[i] = (int*)malloc(nb * sizeof(int)); }
    for (int i = 0; i < nb; i++)
        for (int j = i; j < nb; j++) {
            int64_t k = dm_index(h, h->Mps[j] - h->Mps[i]);
            if (k >= 0) {
                h->bf_dm[i][h->nbf[i]] = (int)k;
                h->bfs[i][h->nbf[i]++] = j;
                h->bis[j][h->nbis[j]++] = i;
            }
        }
    h->plist = (ifvec**)calloc(nb, sizeof(ifvec*));
    h->nlist = (ifvec**)calloc(nb, sizeof(ifvec*));
    #pragma omp parallel for schedule(dynamic, 1)
    for (int bi = 0; bi < nb; bi++) {
        int64_t tot = 0;
        for (int j = 0; j < h->nbf[bi]; j++) { int k = h->bf_dm[bi][j]; tot += h->pnoff[k + 1] - h->pnoff[k]; }
        h->plist[bi] = (ifvec*)calloc(tot > 0 ? tot : 1, sizeof(ifvec));
        h->nlist[bi] = (ifvec*)calloc(tot > 0 ? tot : 1, sizeof(ifvec));
        int64_t base = 0;
        for (int j = 0; j < h->nbf[bi]; j++) {          /* calc_1b_jumps! */
            int bf = h->bfs[bi][j]; int k = h->bf_dm[bi][j];
            int64_t o0 = h->pnoff[k], no = h->pnoff[k + 1] - o0;
            for (int64_t nth = 0; nth < no; nth++) {
                const int64_t* op = h->pnops + 4 * (o0 + nth);
                if (h->pnV[o0 + nth] == 0.0) continue;
                ifvec* pl = &h->plist[bi][base + nth];
                ifvec* nl = &h->nlist[bi][base + nth];
                int64_t npi = h->poff[bi + 1] - h->poff[bi];
                for (int64_t i = 0; i < npi; i++) {
                    int64_t fin; int neg;
                    if (!hop1(h->pbits[h->poff[bi] + i], op[0], op[2], &fin, &neg)) continue;
                    int64_t r = bisearch_desc(h->pbits + h->poff[bf], h->poff[bf + 1] - h->poff[bf], fin);
                    if (r < 0) continue;
                    ifph e = {(int32_t)i, (int32_t)r, (int8_t)neg};
                    ifpush(pl, e);
                }
                int64_t nni = h->noff[bi + 1] - h->noff[bi];
                for (int64_t i = 0; i < nni; i++) {
                    int64_t fin; int neg;
                    if (!hop1(h->nbits[h->noff[bi] + i], op[1], op[3], &fin, &neg)) continue;
                    int64_t r = bisearch_desc(h->nbits + h->noff[bf], h->noff[bf + 1] - h->noff[bf], fin);
                    if (r < 0) continue;
                    ifph e = {(int32_t)i, (int32_t)r, (int8_t)neg};
                    ifpush(nl, e);
                }
            }
            base += no;
        }
    }
    return h;
}

void orc_free(void* hv) {
    orc_t* h = (orc_t*)hv;
    if (!h) return;
    for (int s = 0; s < 2; s++) {
        int64_t n = s == 0 ? h->poff[h->nb] : h->noff[h->nb];
        for (int64_t g = 0; g < n; g++) free(h->t1[s][g].a);
        free(h->t1[s]);
        free(h->ops[s]); free(h->V[s]);
    }
    for (int bi = 0; bi < h->nb; bi++) {
        int64_t tot = 0;
        for (int j = 0; j < h->nbf[bi]; j++) { int k = h->bf_dm[bi][j]; tot += h->pnoff[k + 1] - h->pnoff[k]; }
        for (int64_t q = 0; q < tot; q++) { free(h->plist[bi][q].a); free(h->nlist[bi][q].a); }
        free(h->plist[bi]); free(h->nlist[bi]);
        free(h->bfs[bi]); free(h->bis[bi]); free(h->bf_dm[bi]);
    }
    free(h->plist); free(h->nlist); free(h->bfs); free(h->bis); free(h->bf_dm); free(h->nbf); free(h->nbis);
    free(h->pbits); free(h->poff); free(h->nbits); free(h->noff); free(h->tdims); free(h->Mps);
    free(h->dms); free(h->pnoff); free(h->pnops); free(h->pnV);
    free(h->spe_p); free(h->spe_n); free(h->jp); free(h->jn);
    free(h);
}

/* position of bf in bfs[bi] (bisearch_ord! on the ascending list) */
static int bf_pos(const orc_t* h, int bi, int bf) {
    int lo = 0, hi = h->nbf[bi] - 1;
    while (lo <= hi) { int m = (lo + hi) / 2; if (h->bfs[bi][m] > bf) hi = m - 1; else if (h->bfs[bi][m] < bf) lo = m + 1; else return m; }
    return -1;
}
static int64_t pair_base(const orc_t* h, int bi, int j) {
    int64_t base = 0;
    for (int q = 0; q < j; q++) { int k = h->bf_dm[bi][q]; base += h->pnoff[k + 1] - h->pnoff[k]; }
    return base;
}

void orc_operate_H(void* hv, const double* wf, double* twf) {
    const orc_t* h = (const orc_t*)hv;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int bi = 0; bi < h->nb; bi++) {
        int64_t idim = h->tdims[bi];
        int64_t lNp = h->poff[bi + 1] - h->poff[bi];
        int64_t lNn = h->noff[bi + 1] - h->noff[bi];
        /* pn1 */
        int64_t base = 0;
        for (int j = 0; j < h->nbf[bi]; j++) {
            int bf = h->bfs[bi][j]; int k = h->bf_dm[bi][j];
            int64_t o0 = h->pnoff[k], no = h->pnoff[k + 1] - o0;
            int64_t fdim = h->tdims[bf], lNnf = h->noff[bf + 1] - h->noff[bf];
            for (int64_t nth = 0; nth < no; nth++) {
                double V = h->pnV[o0 + nth];
                const ifvec* pl = &h->plist[bi][base + nth];
                const ifvec* nl = &h->nlist[bi][base + nth];
                for (int64_t a = 0; a < pl->n; a++) {
                    int64_t tMi = idim + (int64_t)pl->a[a].i * lNn;
                    int64_t tMf = fdim + (int64_t)pl->a[a].f * lNnf;
                    int8_t pp = pl->a[a].ph;
                    for (int64_t b = 0; b < nl->n; b++) {
                        int64_t Mi = tMi + nl->a[b].i, Mf = tMf + nl->a[b].f;
                        twf[Mi] += (pp != nl->a[b].ph ? -V : V) * wf[Mf];
                    }
                }
            }
            base += no;
        }
        /* pn2 */
        for (int j = 0; j < h->nbis[bi] - 1; j++) {
            int tbi = h->bis[bi][j];
            int jj = bf_pos(h, tbi, bi);
            int k = h->bf_dm[tbi][jj];
            int64_t pb = pair_base(h, tbi, jj);
            int64_t o0 = h->pnoff[k], no = h->pnoff[k + 1] - o0;
            int64_t fdim = h->tdims[tbi], lNni = h->noff[tbi + 1] - h->noff[tbi];
            for (int64_t nth = 0; nth < no; nth++) {
                double V = h->pnV[o0 + nth];
                const ifvec* pl = &h->plist[tbi][pb + nth];
                const ifvec* nl = &h->nlist[tbi][pb + nth];
                for (int64_t a = 0; a < pl->n; a++) {
                    int64_t tMi = fdim + (int64_t)pl->a[a].i * lNni;
                    int64_t tMf = idim + (int64_t)pl->a[a].f * lNn;
                    int8_t pp = pl->a[a].ph;
                    for (int64_t b = 0; b < nl->n; b++) {
                        int64_t Mi = tMi + nl->a[b].i, Mf = tMf + nl->a[b].f;
                        twf[Mf] += (pp != nl->a[b].ph ? -V : V) * wf[Mi];
                    }
                }
            }
        }
        /* pp */
        for (int64_t i = 0; i < lNp; i++) {
            const t1vec* tv = &h->t1[0][h->poff[bi] + i];
            int64_t tMi = idim + i * lNn;
            for (int64_t e = 0; e < tv->n; e++) {
                int64_t tMf = idim + (int64_t)tv->a[e].f * lNn;
                double fac = tv->a[e].coef;
                for (int64_t n = 0; n < lNn; n++) {
                    twf[tMf + n] += fac * wf[tMi + n];
                    twf[tMi + n] += fac * wf[tMf + n];
                }
            }
        }
        /* nn */
        for (int64_t i = 0; i < lNn; i++) {
            const t1vec* tv = &h->t1[1][h->noff[bi] + i];
            for (int64_t e = 0; e < tv->n; e++) {
                int64_t f = tv->a[e].f;
                double fac = tv->a[e].coef;
                for (int64_t p = 0; p < lNp; p++) {
                    int64_t Mi = idim + p * lNn + i, Mf = idim + p * lNn + f;
                    twf[Mf] += fac * wf[Mi];
                    twf[Mi] += fac * wf[Mf];
                }
            }
        }
        /* one-body */
        for (int64_t p = 0; p < lNp; p++) {
            const double* jpp = h->jp + (h->poff[bi] + p) * h->norb_p;
            double ep = 0.0;
            for (int o = 0; o < h->norb_p; o++) ep += h->spe_p[o] * jpp[o];
            for (int64_t n = 0; n < lNn; n++) {
                const double* jnn = h->jn + (h->noff[bi] + n) * h->norb_n;
                double en = 0.0;
                for (int o = 0; o < h->norb_n; o++) en += h->spe_n[o] * jnn[o];
                int64_t Mi = idim + p * lNn + n;
                twf[Mi] += (ep + en) * wf[Mi];
            }
        }
    }
}

/* counts: [pp entries, nn entries, (pair,op) with both lists non-empty, p-hop entries,
 *          n-hop entries, W (FMAs per operate_H!), mdim, pairs] */
void orc_counts(void* hv, int64_t* out) {
    const orc_t* h = (const orc_t*)hv;
    memset(out, 0, 8 * sizeof(int64_t));
    int64_t W = 0;
    for (int s = 0; s < 2; s++) {
        const int64_t* off = s == 0 ? h->poff : h->noff;
        for (int b = 0; b < h->nb; b++) {
            int64_t other = s == 0 ? (h->noff[b + 1] - h->noff[b]) : (h->poff[b + 1] - h->poff[b]);
            for (int64_t g = off[b]; g < off[b + 1]; g++) { out[s] += h->t1[s][g].n; W += 2 * h->t1[s][g].n * other; }
        }
    }
    for (int bi = 0; bi < h->nb; bi++) {
        int64_t base = 0;
        for (int j = 0; j < h->nbf[bi]; j++) {
            int k = h->bf_dm[bi][j]; int64_t no = h->pnoff[k + 1] - h->pnoff[k];
            int bf = h->bfs[bi][j];
            for (int64_t q = 0; q < no; q++) {
                int64_t np = h->plist[bi][base + q].n, nn = h->nlist[bi][base + q].n;
                out[3] += np; out[4] += nn;
                if (np && nn) out[2]++;
                W += np * nn * (bf == bi ? 1 : 2);
            }
            base += no;
            out[7]++;
        }
    }
    W += h->tdims[h->nb];
    out[5] = W; out[6] = h->tdims[h->nb];
}

/* every T1info entry: block, i, f, coef (reference order) */
void orc_export_pp(void* hv, int s, int64_t* blk, int64_t* ii, int64_t* ff, double* cf) {
    const orc_t* h = (const orc_t*)hv;
    const int64_t* off = s == 0 ? h->poff : h->noff;
    int64_t q = 0;
    for (int b = 0; b < h->nb; b++)
        for (int64_t g = off[b]; g < off[b + 1]; g++)
            for (int64_t e = 0; e < h->t1[s][g].n; e++) {
                blk[q] = b; ii[q] = g - off[b]; ff[q] = h->t1[s][g].a[e].f; cf[q] = h->t1[s][g].a[e].coef; q++;
            }
}

/* every pn hop entry of species s: bi, bf, global op index, i, f, phase bit */
void orc_export_pn(void* hv, int s, int64_t* bi_o, int64_t* bf_o, int64_t* op_o, int64_t* i_o, int64_t* f_o, int8_t* ph_o) {
    const orc_t* h = (const orc_t*)hv;
    int64_t q = 0;
    for (int bi = 0; bi < h->nb; bi++) {
        int64_t base = 0;
        for (int j = 0; j < h->nbf[bi]; j++) {
            int k = h->bf_dm[bi][j]; int64_t o0 = h->pnoff[k], no = h->pnoff[k + 1] - o0;
            for (int64_t nth = 0; nth < no; nth++) {
                const ifvec* L = s == 0 ? &h->plist[bi][base + nth] : &h->nlist[bi][base + nth];
                for (int64_t e = 0; e < L->n; e++) {
                    bi_o[q] = bi; bf_o[q] = h->bfs[bi][j]; op_o[q] = o0 + nth;
                    i_o[q] = L->a[e].i; f_o[q] = L->a[e].f; ph_o[q] = L->a[e].ph; q++;
                }
            }
            base += no;
        }
    }
}
