// H·v on gfx950: one workgroup per output proton row.
//
// Output element (bf, fp, fn) (block, proton SD, neutron SD; global index
// tdims[bf] + fp*nN + fn, the reference's order) is owned by one lane; all
// contributions are gathered, so no atomics are needed and the result does not
// depend on the launch shape:
//   diagonal : (e_p[fp] + e_n[fn]) v[fp,fn]
//   nn       : sum_k c_k v[fp, col_k]         (row fp of v staged in LDS)
//   pp       : sum_e c_e v[ip_e, fn]          (coalesced rows of the same block)
//   pn       : for every input block bi coupled by a proton dM, for every proton
//              hop ip -> fp (operator ac, sign s_p):
//                sum_{neutron hops in -> fn (operator bd, sign s_n)}
//                    s_p s_n V_eff[dM][ac][bd] v[bi: ip, in]
//              The input rows s_p v[ip,:] and the V_eff rows of the hop's
//              operators are staged in LDS; each lane walks its own neutron hop
//              list (sliced ELL: 64 consecutive fn per slice, coalesced).
// The reference's loop nest for the same sums is operate_H!
// (src/ShellModel/lanczos_methods.jl:153-248).
#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "sm_device.h"

namespace smgpu {

DBuf::~DBuf() {
    if (p) (void)hipFree(p);
}
void DBuf::alloc(size_t n) {
    if (p) { (void)hipFree(p); p = nullptr; }
    bytes = n;
    if (n) SM_HIP(hipMalloc(&p, n));
}

DeviceHam::~DeviceHam() {
    if (stream) (void)hipStreamDestroy(stream);
}

template <typename T, int NPT>
__global__ __launch_bounds__(256) void k_hv(HvTables A, const T* __restrict__ v, T* __restrict__ w) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* xs_self = reinterpret_cast<T*>(smem_raw);
    T* xrows = xs_self + A.nN_max;
    T* vrows = xrows + (size_t)A.pch * A.nN_max;
    const int r = blockIdx.x;
    const int tid = threadIdx.x;
    const int bf = A.row_block[r];
    const int fp = r - A.block_row0[bf];
    const int nN = A.nN[bf];
    const int64_t bstart = A.tdims[bf];
    const int64_t base = bstart + (int64_t)fp * nN;
    const T* ep = reinterpret_cast<const T*>(A.ep);
    const T* en = reinterpret_cast<const T*>(A.en) + A.en_off[bf];
    for (int i = tid; i < nN; i += 256) xs_self[i] = v[base + i];
    __syncthreads();

    T acc[NPT];
    const T epr = ep[r];
    const T* nnv = reinterpret_cast<const T*>(A.nn_val);
    const int64_t nsb = A.nn_slice_base[bf];
#pragma unroll
    for (int k = 0; k < NPT; k++) {
        const int fn = tid + k * 256;
        acc[k] = T(0);
        if (fn < nN) {
            T a = (epr + en[fn]) * xs_self[fn];
            const int64_t s = nsb + (fn >> 6);
            const int lane = fn & 63;
            const int64_t off = A.nn_slice_off[s] + lane;
            const int len = A.nn_slice_len[s];
            for (int q = 0; q < len; q++) {
                const int64_t e = off + (int64_t)q * 64;
                a += nnv[e] * xs_self[A.nn_col[e]];
            }
            acc[k] = a;
        }
    }
    // pp: rows of the same block
    const T* ppv = reinterpret_cast<const T*>(A.pp_val);
    const int64_t pe1 = A.pp_rowptr[r + 1];
    for (int64_t e = A.pp_rowptr[r]; e < pe1; e++) {
        const T c = ppv[e];
        const T* src = v + bstart + (int64_t)A.pp_col[e] * nN;
#pragma unroll
        for (int k = 0; k < NPT; k++) {
            const int fn = tid + k * 256;
            if (fn < nN) acc[k] += c * src[fn];
        }
    }
    // pn
    const T* veff = reinterpret_cast<const T*>(A.veff);
    const int q1 = A.pair_ptr[bf + 1];
    for (int q = A.pair_ptr[bf]; q < q1; q++) {
        const int pt = A.pair_ptab[q];
        const int32_t* prp = A.ptab_rowptr + A.ptab_rowptr_base[pt];
        const int p0 = prp[fp], p1 = prp[fp + 1];
        if (p0 == p1) continue;
        const int bi = A.pair_bi[q];
        const int dm = A.pair_dm[q];
        const int nt = A.pair_ntab[q];
        const uint32_t* pent = A.ptab_ent + A.ptab_ent_base[pt];
        const int nNi = A.nN[bi];
        const int64_t ibase = A.tdims[bi];
        const T* V2 = veff + A.veff_off[dm];
        const int c2 = A.veff_cols2[dm];
        const int64_t sb = A.ntab_slice_base[nt];
        for (int pc = p0; pc < p1; pc += A.pch) {
            const int np = min(A.pch, p1 - pc);
            __syncthreads();
            for (int j = 0; j < np; j++) {
                const uint32_t e = pent[pc + j];
                const int ip = (int)(e & 0xFFFFFu);
                const int op = (int)((e >> 20) & 0x7FFu);
                const T sg = (e >> 31) ? T(-1) : T(1);
                const T* src = v + ibase + (int64_t)ip * nNi;
                T* dst = xrows + (size_t)j * A.nN_max;
                for (int i = tid; i < nNi; i += 256) dst[i] = sg * src[i];
                const T* vr = V2 + (int64_t)op * c2;
                T* vd = vrows + (size_t)j * A.cols2_max;
                for (int i = tid; i < c2; i += 256) vd[i] = vr[i];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < NPT; k++) {
                const int fn = tid + k * 256;
                if (fn < nN) {
                    const int64_t s = sb + (fn >> 6);
                    const int lane = fn & 63;
                    const int64_t off = A.ntab_slice_off[s] + lane;
                    const int len = A.ntab_slice_len[s];
                    T a = acc[k];
                    for (int qq = 0; qq < len; qq++) {
                        const uint32_t e = A.ntab_ent[off + (int64_t)qq * 64];
                        const int in = (int)(e & 0xFFFFFu);
                        const int col = (int)(e >> 20);
                        for (int j = 0; j < np; j++)
                            a += vrows[(size_t)j * A.cols2_max + col] * xrows[(size_t)j * A.nN_max + in];
                    }
                    acc[k] = a;
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NPT; k++) {
        const int fn = tid + k * 256;
        if (fn < nN) w[base + fn] = acc[k];
    }
}

template <typename T>
static void launch_t(const HvTables& A, const T* v, T* w, hipStream_t s) {
    dim3 grid(A.R), block(256);
    switch (A.npt) {
        case 1: hipLaunchKernelGGL((k_hv<T, 1>), grid, block, A.lds_bytes, s, A, v, w); break;
        case 2: hipLaunchKernelGGL((k_hv<T, 2>), grid, block, A.lds_bytes, s, A, v, w); break;
        case 4: hipLaunchKernelGGL((k_hv<T, 4>), grid, block, A.lds_bytes, s, A, v, w); break;
        case 8: hipLaunchKernelGGL((k_hv<T, 8>), grid, block, A.lds_bytes, s, A, v, w); break;
        case 16: hipLaunchKernelGGL((k_hv<T, 16>), grid, block, A.lds_bytes, s, A, v, w); break;
        default: throw std::runtime_error("neutron block wider than 4096 determinants");
    }
    SM_HIP(hipGetLastError());
}

void launch_hv(const DeviceHam& h, int prec, const void* v, void* w, hipStream_t s) {
    if (h.R == 0) return;
    if (prec == 0) launch_t<double>(h.view[0], (const double*)v, (double*)w, s);
    else launch_t<float>(h.view[1], (const float*)v, (float*)w, s);
}

// ---------------------------------------------------------------- upload
namespace {

template <typename T> std::vector<T> cast_vec(const std::vector<double>& v) {
    std::vector<T> o(v.size());
    for (size_t i = 0; i < v.size(); i++) o[i] = (T)v[i];
    return o;
}

struct HostLayout {
    std::vector<int32_t> row_block, block_row0, nP, nN;
    std::vector<int64_t> tdims, en_off;
    std::vector<double> ep, en;
    std::vector<int64_t> pp_rowptr;
    std::vector<int32_t> pp_col;
    std::vector<double> pp_val;
    std::vector<int64_t> nn_slice_base, nn_slice_off;
    std::vector<int32_t> nn_slice_len, nn_col;
    std::vector<double> nn_val;
    std::vector<int32_t> pair_ptr, pair_bi, pair_dm, pair_ptab, pair_ntab;
    std::vector<int64_t> ptab_rowptr_base, ptab_ent_base;
    std::vector<int32_t> ptab_rowptr;
    std::vector<uint32_t> ptab_ent;
    std::vector<int64_t> ntab_slice_base, ntab_slice_off;
    std::vector<int32_t> ntab_slice_len;
    std::vector<uint32_t> ntab_ent;
    std::vector<int64_t> veff_off;
    std::vector<int32_t> veff_cols2;
    std::vector<double> veff;
    int nN_max = 0, cols2_max = 0;
    int64_t pn_terms = 0, pn_terms_exact = 0, pp_terms = 0, nn_terms = 0;
};

void coef_layout(const Basis& b, const Tables& t, HostLayout& L) {
    L.ep.clear(); L.en.clear(); L.pp_val.clear(); L.nn_val.clear(); L.veff.clear();
    for (int k = 0; k < b.nb; k++) {
        const auto& e = t.ep[b.pset[k]];
        L.ep.insert(L.ep.end(), e.begin(), e.end());
        const auto& f = t.en[b.nset[k]];
        L.en.insert(L.en.end(), f.begin(), f.end());
        const PairSym& m = t.pp[b.pset[k]];
        L.pp_val.insert(L.pp_val.end(), m.val.begin(), m.val.end());
        const PairSym& nm = t.nn[b.nset[k]];
        int nN = (int)b.sp[1].sets[b.nset[k]].size();
        for (int s0 = 0; s0 < nN; s0 += 64) {
            int len = 0;
            for (int f = s0; f < std::min(nN, s0 + 64); f++) len = std::max<int>(len, (int)(nm.rowptr[f + 1] - nm.rowptr[f]));
            for (int q = 0; q < len; q++)
                for (int lane = 0; lane < 64; lane++) {
                    int f = s0 + lane;
                    double val = 0.0;
                    if (f < nN && nm.rowptr[f] + q < nm.rowptr[f + 1]) val = nm.val[nm.rowptr[f] + q];
                    L.nn_val.push_back(val);
                }
        }
    }
    for (size_t k = 0; k < t.veff.size(); k++) {
        int rows = t.veff_rows[k], cols = t.veff_cols[k];
        int c2 = 2 * (cols + 1);
        for (int r = 0; r < rows; r++) {
            for (int c = 0; c < cols; c++) {
                double v = t.veff[k][(size_t)r * cols + c];
                L.veff.push_back(v);
                L.veff.push_back(-v);
            }
            L.veff.push_back(0.0);
            L.veff.push_back(0.0);
        }
        (void)c2;
    }
}

void index_layout(const Basis& b, const Tables& t, HostLayout& L) {
    const Species& P = b.sp[0];
    const Species& N = b.sp[1];
    L.tdims = b.tdims;
    int R = 0;
    for (int k = 0; k < b.nb; k++) {
        int np = (int)P.sets[b.pset[k]].size(), nn = (int)N.sets[b.nset[k]].size();
        L.block_row0.push_back(R);
        L.nP.push_back(np);
        L.nN.push_back(nn);
        L.nN_max = std::max(L.nN_max, nn);
        for (int i = 0; i < np; i++) L.row_block.push_back(k);
        R += np;
    }
    // pp CSR per global row
    L.pp_rowptr.push_back(0);
    for (int k = 0; k < b.nb; k++) {
        const PairSym& m = t.pp[b.pset[k]];
        int np = L.nP[k];
        for (int i = 0; i < np; i++) {
            for (int64_t e = m.rowptr[i]; e < m.rowptr[i + 1]; e++) L.pp_col.push_back(m.col[e]);
            L.pp_rowptr.push_back((int64_t)L.pp_col.size());
            L.pp_terms += (m.rowptr[i + 1] - m.rowptr[i]) * (int64_t)L.nN[k];
        }
    }
    // nn sliced ELL per block
    int64_t en_acc = 0;
    for (int k = 0; k < b.nb; k++) {
        L.en_off.push_back(en_acc);
        en_acc += L.nN[k];
        const PairSym& m = t.nn[b.nset[k]];
        int nN = L.nN[k];
        L.nn_slice_base.push_back((int64_t)L.nn_slice_off.size());
        for (int s0 = 0; s0 < nN; s0 += 64) {
            int len = 0;
            for (int f = s0; f < std::min(nN, s0 + 64); f++) len = std::max<int>(len, (int)(m.rowptr[f + 1] - m.rowptr[f]));
            L.nn_slice_off.push_back((int64_t)L.nn_col.size());
            L.nn_slice_len.push_back(len);
            for (int q = 0; q < len; q++)
                for (int lane = 0; lane < 64; lane++) {
                    int f = s0 + lane;
                    int col = 0;
                    if (f < nN && m.rowptr[f] + q < m.rowptr[f + 1]) col = m.col[m.rowptr[f] + q];
                    L.nn_col.push_back(col);
                }
        }
        for (int f = 0; f < nN; f++) L.nn_terms += (m.rowptr[f + 1] - m.rowptr[f]) * (int64_t)L.nP[k];
    }
    // hop tables: proton tables keep CSR, neutron tables become sliced ELL
    std::vector<int> pt_of(t.hops.size(), -1), nt_of(t.hops.size(), -1);
    for (size_t h = 0; h < t.hops.size(); h++) {
        const HopTable& H = t.hops[h];
        if (H.species == 0) {
            pt_of[h] = (int)L.ptab_rowptr_base.size();
            L.ptab_rowptr_base.push_back((int64_t)L.ptab_rowptr.size());
            L.ptab_ent_base.push_back((int64_t)L.ptab_ent.size());
            for (int64_t x : H.rowptr) L.ptab_rowptr.push_back((int32_t)x);
            L.ptab_ent.insert(L.ptab_ent.end(), H.ent.begin(), H.ent.end());
        } else {
            nt_of[h] = (int)L.ntab_slice_base.size();
            L.ntab_slice_base.push_back((int64_t)L.ntab_slice_off.size());
            int kp = t.ob[0].index_of_dM(-H.dM);
            int cols = kp >= 0 ? t.veff_cols[kp] : 0;
            int nf = (int)H.rowptr.size() - 1;
            for (int s0 = 0; s0 < nf; s0 += 64) {
                int len = 0;
                for (int f = s0; f < std::min(nf, s0 + 64); f++) len = std::max<int>(len, (int)(H.rowptr[f + 1] - H.rowptr[f]));
                L.ntab_slice_off.push_back((int64_t)L.ntab_ent.size());
                L.ntab_slice_len.push_back(len);
                for (int q = 0; q < len; q++)
                    for (int lane = 0; lane < 64; lane++) {
                        int f = s0 + lane;
                        uint32_t ent = (uint32_t)(2 * cols) << 20;         // zero column
                        if (f < nf && H.rowptr[f] + q < H.rowptr[f + 1]) {
                            uint32_t e = H.ent[H.rowptr[f] + q];
                            uint32_t in = e & 0xFFFFFu, op = (e >> 20) & 0x7FFu, neg = e >> 31;
                            ent = in | ((2 * op + neg) << 20);
                        }
                        L.ntab_ent.push_back(ent);
                    }
            }
        }
    }
    // V_eff offsets
    int64_t voff = 0;
    for (size_t k = 0; k < t.veff.size(); k++) {
        int c2 = 2 * (t.veff_cols[k] + 1);
        if (c2 >= 4096) throw std::runtime_error("too many neutron one-body operators for the packed entry");
        L.veff_off.push_back(voff);
        L.veff_cols2.push_back(c2);
        L.cols2_max = std::max(L.cols2_max, c2);
        voff += (int64_t)t.veff_rows[k] * c2;
    }
    // block pairs (output bf, input bi)
    L.pair_ptr.push_back(0);
    for (int bf = 0; bf < b.nb; bf++) {
        for (int bi = 0; bi < b.nb; bi++) {
            int dM = b.Mp[bf] - b.Mp[bi];
            int k = t.ob[0].index_of_dM(dM);
            if (k < 0 || t.ob[1].index_of_dM(-dM) < 0) continue;
            auto ip = t.hop_index.find({0, b.pset[bi], b.pset[bf]});
            auto in = t.hop_index.find({1, b.nset[bi], b.nset[bf]});
            if (ip == t.hop_index.end() || in == t.hop_index.end()) continue;
            L.pair_bi.push_back(bi);
            L.pair_dm.push_back(k);
            L.pair_ptab.push_back(pt_of[ip->second]);
            L.pair_ntab.push_back(nt_of[in->second]);
            const HopTable& HP = t.hops[ip->second];
            const HopTable& HN = t.hops[in->second];
            int64_t nsum = 0;
            for (size_t f = 0; f + 1 < HN.rowptr.size(); f++) nsum += HN.rowptr[f + 1] - HN.rowptr[f];
            int64_t npad = 0;
            int nf = (int)HN.rowptr.size() - 1;
            for (int s0 = 0; s0 < nf; s0 += 64) {
                int len = 0;
                for (int f = s0; f < std::min(nf, s0 + 64); f++) len = std::max<int>(len, (int)(HN.rowptr[f + 1] - HN.rowptr[f]));
                npad += (int64_t)len * std::min(64, nf - s0);
            }
            int64_t psum = (int64_t)HP.ent.size();
            L.pn_terms_exact += psum * nsum;
            L.pn_terms += psum * npad;
        }
        L.pair_ptr.push_back((int32_t)L.pair_bi.size());
    }
}

template <typename T> void set_view(DeviceHam& h, int prec, const HostLayout& L) {
    HvTables& A = h.view[prec];
    A.nb = h.nb; A.R = h.R;
    A.row_block = h.row_block.as<int32_t>(); A.block_row0 = h.block_row0.as<int32_t>();
    A.nP = h.nP.as<int32_t>(); A.nN = h.nN.as<int32_t>(); A.tdims = h.tdims.as<int64_t>();
    A.ep = h.ep[prec].p; A.en_off = h.en_off.as<int64_t>(); A.en = h.en[prec].p;
    A.pp_rowptr = h.pp_rowptr.as<int64_t>(); A.pp_col = h.pp_col.as<int32_t>(); A.pp_val = h.pp_val[prec].p;
    A.nn_slice_base = h.nn_slice_base.as<int64_t>(); A.nn_slice_off = h.nn_slice_off.as<int64_t>();
    A.nn_slice_len = h.nn_slice_len.as<int32_t>(); A.nn_col = h.nn_col.as<int32_t>(); A.nn_val = h.nn_val[prec].p;
    A.pair_ptr = h.pair_ptr.as<int32_t>(); A.pair_bi = h.pair_bi.as<int32_t>(); A.pair_dm = h.pair_dm.as<int32_t>();
    A.pair_ptab = h.pair_ptab.as<int32_t>(); A.pair_ntab = h.pair_ntab.as<int32_t>();
    A.ptab_rowptr_base = h.ptab_rowptr_base.as<int64_t>(); A.ptab_rowptr = h.ptab_rowptr.as<int32_t>();
    A.ptab_ent_base = h.ptab_ent_base.as<int64_t>(); A.ptab_ent = h.ptab_ent.as<uint32_t>();
    A.ntab_slice_base = h.ntab_slice_base.as<int64_t>(); A.ntab_slice_off = h.ntab_slice_off.as<int64_t>();
    A.ntab_slice_len = h.ntab_slice_len.as<int32_t>(); A.ntab_ent = h.ntab_ent.as<uint32_t>();
    A.veff_off = h.veff_off.as<int64_t>(); A.veff_cols2 = h.veff_cols2.as<int32_t>(); A.veff = h.veff[prec].p;
    A.nN_max = std::max(1, L.nN_max);
    A.cols2_max = std::max(2, L.cols2_max);
    const size_t budget = 64 * 1024 / sizeof(T);
    if ((size_t)A.nN_max + A.nN_max + A.cols2_max > budget)
        throw std::runtime_error("neutron block too wide for the LDS budget");
    int pch = (int)((budget - A.nN_max) / (size_t)(A.nN_max + A.cols2_max));
    A.pch = std::max(1, std::min(16, pch));
    A.lds_bytes = ((size_t)A.nN_max + (size_t)A.pch * (A.nN_max + A.cols2_max)) * sizeof(T);
    int need = (A.nN_max + 255) / 256;
    int npt = 1;
    while (npt < need) npt *= 2;
    A.npt = npt;
}

template <typename T> void upload_coef(DeviceHam& h, int prec, const HostLayout& L) {
    h.ep[prec].upload(cast_vec<T>(L.ep));
    h.en[prec].upload(cast_vec<T>(L.en));
    h.pp_val[prec].upload(cast_vec<T>(L.pp_val));
    h.nn_val[prec].upload(cast_vec<T>(L.nn_val));
    h.veff[prec].upload(cast_vec<T>(L.veff));
}

}  // namespace

std::unique_ptr<DeviceHam> upload_ham(const Basis& b, const Tables& t, int device) {
    SM_HIP(hipSetDevice(device));
    auto h = std::make_unique<DeviceHam>();
    h->device = device;
    SM_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HostLayout L;
    index_layout(b, t, L);
    coef_layout(b, t, L);
    h->mdim = b.mdim;
    h->nb = b.nb;
    h->R = (int)L.row_block.size();
    h->row_block.upload(L.row_block); h->block_row0.upload(L.block_row0);
    h->nP.upload(L.nP); h->nN.upload(L.nN); h->tdims.upload(L.tdims); h->en_off.upload(L.en_off);
    h->pp_rowptr.upload(L.pp_rowptr); h->pp_col.upload(L.pp_col);
    h->nn_slice_base.upload(L.nn_slice_base); h->nn_slice_off.upload(L.nn_slice_off);
    h->nn_slice_len.upload(L.nn_slice_len); h->nn_col.upload(L.nn_col);
    h->pair_ptr.upload(L.pair_ptr); h->pair_bi.upload(L.pair_bi); h->pair_dm.upload(L.pair_dm);
    h->pair_ptab.upload(L.pair_ptab); h->pair_ntab.upload(L.pair_ntab);
    h->ptab_rowptr_base.upload(L.ptab_rowptr_base); h->ptab_rowptr.upload(L.ptab_rowptr);
    h->ptab_ent_base.upload(L.ptab_ent_base); h->ptab_ent.upload(L.ptab_ent);
    h->ntab_slice_base.upload(L.ntab_slice_base); h->ntab_slice_off.upload(L.ntab_slice_off);
    h->ntab_slice_len.upload(L.ntab_slice_len); h->ntab_ent.upload(L.ntab_ent);
    h->veff_off.upload(L.veff_off); h->veff_cols2.upload(L.veff_cols2);
    upload_coef<double>(*h, 0, L);
    upload_coef<float>(*h, 1, L);
    set_view<double>(*h, 0, L);
    set_view<float>(*h, 1, L);
    h->pn_terms = L.pn_terms;
    h->pn_terms_exact = L.pn_terms_exact;
    h->pp_terms = L.pp_terms;
    h->nn_terms = L.nn_terms;
    return h;
}

void update_coefficients(DeviceHam& h, const Basis& b, const Tables& t) {
    SM_HIP(hipSetDevice(h.device));
    HostLayout L;
    coef_layout(b, t, L);
    upload_coef<double>(h, 0, L);
    upload_coef<float>(h, 1, L);
    for (int prec = 0; prec < 2; prec++) {
        HvTables& A = h.view[prec];
        A.ep = h.ep[prec].p; A.en = h.en[prec].p; A.pp_val = h.pp_val[prec].p;
        A.nn_val = h.nn_val[prec].p; A.veff = h.veff[prec].p;
    }
}

std::vector<std::pair<std::string, int64_t>> DeviceHam::table_bytes(int prec) const {
    const size_t s = prec == 0 ? 8 : 4;
    std::vector<std::pair<std::string, int64_t>> v = {
        {"row_block", (int64_t)row_block.bytes}, {"block_row0", (int64_t)block_row0.bytes},
        {"nP", (int64_t)nP.bytes}, {"nN", (int64_t)nN.bytes}, {"tdims", (int64_t)tdims.bytes},
        {"en_off", (int64_t)en_off.bytes},
        {"ep", (int64_t)(ep[prec].bytes)}, {"en", (int64_t)(en[prec].bytes)},
        {"pp_rowptr", (int64_t)pp_rowptr.bytes}, {"pp_col", (int64_t)pp_col.bytes},
        {"pp_val", (int64_t)pp_val[prec].bytes},
        {"nn_slice_base", (int64_t)nn_slice_base.bytes}, {"nn_slice_off", (int64_t)nn_slice_off.bytes},
        {"nn_slice_len", (int64_t)nn_slice_len.bytes}, {"nn_col", (int64_t)nn_col.bytes},
        {"nn_val", (int64_t)nn_val[prec].bytes},
        {"pair_ptr", (int64_t)pair_ptr.bytes}, {"pair_bi", (int64_t)pair_bi.bytes},
        {"pair_dm", (int64_t)pair_dm.bytes}, {"pair_ptab", (int64_t)pair_ptab.bytes},
        {"pair_ntab", (int64_t)pair_ntab.bytes},
        {"ptab_rowptr_base", (int64_t)ptab_rowptr_base.bytes}, {"ptab_rowptr", (int64_t)ptab_rowptr.bytes},
        {"ptab_ent_base", (int64_t)ptab_ent_base.bytes}, {"ptab_ent", (int64_t)ptab_ent.bytes},
        {"ntab_slice_base", (int64_t)ntab_slice_base.bytes}, {"ntab_slice_off", (int64_t)ntab_slice_off.bytes},
        {"ntab_slice_len", (int64_t)ntab_slice_len.bytes}, {"ntab_ent", (int64_t)ntab_ent.bytes},
        {"veff_off", (int64_t)veff_off.bytes}, {"veff_cols2", (int64_t)veff_cols2.bytes},
        {"veff", (int64_t)veff[prec].bytes},
        {"vector_in", (int64_t)(mdim * s)}, {"vector_out", (int64_t)(mdim * s)}};
    return v;
}

}  // namespace smgpu
