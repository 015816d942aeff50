// Magnetic dipole and electric quadrupole moments of Ritz vectors (main_sm's
// calc_moment; transit.jl: prep_obtr_diag :82-245, op_obtr_diag :247-283,
// eval_moment :702-765) and the M1 / E2 transitions between two sets of them
// (transit_main :561-700, prep_obtr_nondiag :286-445, op_obtr_nondiag :447-499).
// The moment operators are the M = 0 components of one-body tensors, so they keep
// every SD in its M-set: w[b, ip, n] gathers the proton matrix over the rows ip'
// of block b and the neutron matrix over its columns n'.  The transition
// operators (component mu) map a right basis at 2M_r into a left one at 2M_l =
// 2M_r + 2mu: left block (Mp, Mn) gathers its proton term from the right block
// (Mp - 2mu, Mn) and its neutron term from (Mp, Mn - 2mu).  One workgroup per
// (left) proton row, lanes along the neutron index, every output element written
// once (deterministic); FP64 throughout.
#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <map>
#include <stdexcept>
#include <vector>

#include "sm_device.h"

namespace smgpu {

struct MomView {
    const int32_t* row_block;
    const int32_t* block_row0;
    const int32_t* nN;
    const int64_t* tdims;
    const int32_t* pset;
    const int32_t* nset;
    const int64_t* rp_base[2];     // per species: first row pointer of each M-set
    const int64_t* ent_base[2];    // per species: first entry of each M-set
    const int64_t* rowptr[2];
    const int32_t* col[2];
    const double* val[2];
};

static constexpr int MTB = 256;

__global__ __launch_bounds__(MTB) void k_moment(MomView A, const double* __restrict__ v, double* __restrict__ w) {
    const int row = blockIdx.x;
    const int b = A.row_block[row];
    const int ip = row - A.block_row0[b];
    const int nN = A.nN[b];
    const int64_t base = A.tdims[b];
    const int ps = A.pset[b], ns = A.nset[b];
    const int64_t* prp = A.rowptr[0] + A.rp_base[0][ps];
    const int64_t pe0 = A.ent_base[0][ps] + prp[ip], pe1 = A.ent_base[0][ps] + prp[ip + 1];
    const int64_t* nrp = A.rowptr[1] + A.rp_base[1][ns];
    const int64_t nb0 = A.ent_base[1][ns];
    const int64_t orow = base + (int64_t)ip * nN;
    for (int n = threadIdx.x; n < nN; n += MTB) {
        double acc = 0.0;
        for (int64_t e = pe0; e < pe1; e++) acc += A.val[0][e] * v[base + (int64_t)A.col[0][e] * nN + n];
        for (int64_t e = nb0 + nrp[n]; e < nb0 + nrp[n + 1]; e++) acc += A.val[1][e] * v[orow + A.col[1][e]];
        w[orow + n] = acc;
    }
}

// partial sums of x.y per block (summed in block order on the host)
__global__ __launch_bounds__(MTB) void k_dot_part(const double* __restrict__ x, const double* __restrict__ y, int64_t n,
                                                  double* __restrict__ part) {
    __shared__ double red[MTB];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * MTB + threadIdx.x; i < n; i += (int64_t)gridDim.x * MTB) s += x[i] * y[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = MTB / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

struct TrView {
    const int32_t* row_block;      // left rows
    const int32_t* block_row0;
    const int32_t* nN;             // per left block
    const int64_t* tdims;
    const int32_t* pset;
    const int32_t* nset;
    const int32_t* srcP;           // per left block: right block of the proton term (-1: none)
    const int32_t* srcN;           // per left block: right block of the neutron term (-1: none)
    const int32_t* nN_r;           // per right block
    const int64_t* tdims_r;
    const int64_t* rp_base[2];     // per species and left M-set
    const int64_t* ent_base[2];
    const int64_t* rowptr[2];
    const int32_t* col[2];         // SD index in the source M-set
    const double* val[2];
};

__global__ __launch_bounds__(MTB) void k_transition(TrView A, const double* __restrict__ v, double* __restrict__ w) {
    const int row = blockIdx.x;
    const int b = A.row_block[row];
    const int ip = row - A.block_row0[b];
    const int nN = A.nN[b];
    const int64_t orow = A.tdims[b] + (int64_t)ip * nN;
    const int bp = A.srcP[b], bn = A.srcN[b];
    int64_t pe0 = 0, pe1 = 0, pbase = 0;
    if (bp >= 0) {
        const int ps = A.pset[b];
        const int64_t* prp = A.rowptr[0] + A.rp_base[0][ps];
        pe0 = A.ent_base[0][ps] + prp[ip];
        pe1 = A.ent_base[0][ps] + prp[ip + 1];
        pbase = A.tdims_r[bp];                  // the right block has this block's neutron set (nN columns)
    }
    const int64_t* nrp = nullptr;
    int64_t nb0 = 0, nrow = 0;
    if (bn >= 0) {
        const int ns = A.nset[b];
        nrp = A.rowptr[1] + A.rp_base[1][ns];
        nb0 = A.ent_base[1][ns];
        nrow = A.tdims_r[bn] + (int64_t)ip * A.nN_r[bn];   // the same proton SD in the right block
    }
    for (int n = threadIdx.x; n < nN; n += MTB) {
        double acc = 0.0;
        for (int64_t e = pe0; e < pe1; e++) acc += A.val[0][e] * v[pbase + (int64_t)A.col[0][e] * nN + n];
        if (bn >= 0)
            for (int64_t e = nb0 + nrp[n]; e < nb0 + nrp[n + 1]; e++) acc += A.val[1][e] * v[nrow + A.col[1][e]];
        w[orow + n] = acc;
    }
}

DeviceMoments::~DeviceMoments() {
    if (stream) (void)hipStreamDestroy(stream);
}

std::unique_ptr<DeviceMoments> upload_moments(const MomentOps& mo, const Basis& b, int device) {
    SM_HIP(hipSetDevice(device));
    auto m = std::make_unique<DeviceMoments>();
    m->device = device;
    m->mdim = b.mdim;
    m->nb = b.nb;
    m->twoM = b.twoM;
    m->bpar = mo.bpar;
    std::vector<int32_t> rb, r0, nN, ps(b.pset.begin(), b.pset.end()), ns(b.nset.begin(), b.nset.end());
    for (int k = 0; k < b.nb; k++) {
        const int np = (int)b.sp[0].sets[b.pset[k]].size();
        r0.push_back((int32_t)rb.size());
        nN.push_back((int32_t)b.sp[1].sets[b.nset[k]].size());
        for (int i = 0; i < np; i++) rb.push_back(k);
    }
    m->R = (int)rb.size();
    if (rb.empty()) rb.push_back(0);
    m->row_block.upload(rb);
    m->block_row0.upload(r0);
    m->nN.upload(nN);
    m->tdims.upload(b.tdims);
    m->pset.upload(ps);
    m->nset.upload(ns);
    for (int op = 0; op < 2; op++)
        for (int sp = 0; sp < 2; sp++) {
            std::vector<int64_t> rpb, eb, rp;
            std::vector<int32_t> col;
            std::vector<double> val;
            for (const PairSym& P : mo.op[op][sp]) {
                rpb.push_back((int64_t)rp.size());
                eb.push_back((int64_t)col.size());
                rp.insert(rp.end(), P.rowptr.begin(), P.rowptr.end());
                col.insert(col.end(), P.col.begin(), P.col.end());
                val.insert(val.end(), P.val.begin(), P.val.end());
            }
            if (col.empty()) { col.push_back(0); val.push_back(0.0); }
            m->rp_base[op][sp].upload(rpb);
            m->ent_base[op][sp].upload(eb);
            m->rowptr[op][sp].upload(rp);
            m->col[op][sp].upload(col);
            m->val[op][sp].upload(val);
        }
    SM_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    return m;
}

void moments_apply(const DeviceMoments& m, int which, const double* v, double* w, hipStream_t s) {
    if (which != 0 && which != 1) throw std::invalid_argument("operator: 0 (M1) or 1 (E2)");
    if (m.R == 0) return;
    MomView A;
    A.row_block = m.row_block.as<int32_t>();
    A.block_row0 = m.block_row0.as<int32_t>();
    A.nN = m.nN.as<int32_t>();
    A.tdims = m.tdims.as<int64_t>();
    A.pset = m.pset.as<int32_t>();
    A.nset = m.nset.as<int32_t>();
    for (int sp = 0; sp < 2; sp++) {
        A.rp_base[sp] = m.rp_base[which][sp].as<int64_t>();
        A.ent_base[sp] = m.ent_base[which][sp].as<int64_t>();
        A.rowptr[sp] = m.rowptr[which][sp].as<int64_t>();
        A.col[sp] = m.col[which][sp].as<int32_t>();
        A.val[sp] = m.val[which][sp].as<double>();
    }
    hipLaunchKernelGGL(k_moment, dim3(m.R), dim3(MTB), 0, s, A, v, w);
    SM_HIP(hipGetLastError());
}

static double device_dot(DeviceMoments& m, const double* x, const double* y) {
    const int G = (int)std::min<int64_t>(1024, (m.mdim + MTB - 1) / MTB);
    m.part.ensure((size_t)G * 8);
    hipLaunchKernelGGL(k_dot_part, dim3(G), dim3(MTB), 0, m.stream, x, y, m.mdim, m.part.as<double>());
    SM_HIP(hipGetLastError());
    std::vector<double> h(G);
    SM_HIP(hipMemcpyAsync(h.data(), m.part.p, (size_t)G * 8, hipMemcpyDeviceToHost, m.stream));
    SM_HIP(hipStreamSynchronize(m.stream));
    double s = 0.0;
    for (double x2 : h) s += x2;
    return s;
}

void eval_moment(DeviceMoments& m, int nvec, const double* vecs, const double* totJs, double* mu, double* Q) {
    SM_HIP(hipSetDevice(m.device));
    const double nan = std::numeric_limits<double>::quiet_NaN();
    const double c_mu = std::sqrt(4.0 * M_PI / 3.0), c_Q = std::sqrt(16.0 * M_PI / 5.0);
    const int M2 = m.twoM;
    m.v.ensure((size_t)std::max<int64_t>(1, m.mdim) * 8);
    m.w.ensure((size_t)std::max<int64_t>(1, m.mdim) * 8);
    for (int n = 0; n < nvec; n++) {
        mu[n] = nan;
        Q[n] = nan;
        const int J2 = (int)std::lround(2.0 * totJs[n]);
        if (J2 < 0 || m.mdim == 0) continue;
        SM_HIP(hipMemcpyAsync(m.v.p, vecs + (int64_t)n * m.mdim, (size_t)m.mdim * 8, hipMemcpyHostToDevice, m.stream));
        const int sgn = (((J2 - M2) / 2) & 1) ? -1 : 1;     // (-1)^(J - M)
        // lambda = 1 (mu) and 2 (Q): the expectation at M rescaled to the stretched M = J
        const double sel1 = wigner3j(J2, 2, J2, -M2, 0, M2), sel2 = wigner3j(J2, 4, J2, -M2, 0, M2);
        if (sel1 != 0.0) {
            moments_apply(m, 0, m.v.as<double>(), m.w.as<double>(), m.stream);
            const double e = device_dot(m, m.w.as<double>(), m.v.as<double>());
            mu[n] = sgn * e * (wigner3j(J2, 2, J2, -J2, 0, J2) / sel1) * c_mu;
        }
        if (sel2 != 0.0) {
            moments_apply(m, 1, m.v.as<double>(), m.w.as<double>(), m.stream);
            const double e = device_dot(m, m.w.as<double>(), m.v.as<double>());
            Q[n] = e * (wigner3j(J2, 4, J2, -J2, 0, J2) / sel2) * c_Q * sgn;
        }
        SM_HIP(hipStreamSynchronize(m.stream));
    }
}

DeviceTransitions::~DeviceTransitions() {
    if (stream) (void)hipStreamDestroy(stream);
}

static void upload_csr_sets(const std::vector<PairSym>& sets, DBuf& rp_base, DBuf& ent_base, DBuf& rowptr, DBuf& col,
                            DBuf& val) {
    std::vector<int64_t> rpb, eb, rp;
    std::vector<int32_t> c;
    std::vector<double> v;
    for (const PairSym& P : sets) {
        rpb.push_back((int64_t)rp.size());
        eb.push_back((int64_t)c.size());
        rp.insert(rp.end(), P.rowptr.begin(), P.rowptr.end());
        c.insert(c.end(), P.col.begin(), P.col.end());
        v.insert(v.end(), P.val.begin(), P.val.end());
    }
    if (rpb.empty()) rpb.push_back(0);
    if (eb.empty()) eb.push_back(0);
    if (rp.empty()) rp.push_back(0);
    if (c.empty()) { c.push_back(0); v.push_back(0.0); }
    rp_base.upload(rpb);
    ent_base.upload(eb);
    rowptr.upload(rp);
    col.upload(c);
    val.upload(v);
}

std::unique_ptr<DeviceTransitions> upload_transitions(const TransitionOps& to, const Basis& bl, const Basis& br,
                                                      int device) {
    SM_HIP(hipSetDevice(device));
    auto t = std::make_unique<DeviceTransitions>();
    t->device = device;
    t->mdim_l = bl.mdim;
    t->mdim_r = br.mdim;
    t->twoM_l = bl.twoM;
    t->twoM_r = br.twoM;
    t->mu2 = to.mu2;
    t->bpar = to.bpar;
    std::map<std::pair<int, int>, int> rblock;
    for (int k = 0; k < br.nb; k++) rblock[{br.pset[k], br.nset[k]}] = k;
    std::vector<int32_t> rb, r0, nN, ps(bl.pset.begin(), bl.pset.end()), ns(bl.nset.begin(), bl.nset.end()), sP, sN,
        nNr;
    for (int k = 0; k < bl.nb; k++) {
        const int np = (int)bl.sp[0].sets[bl.pset[k]].size();
        r0.push_back((int32_t)rb.size());
        nN.push_back((int32_t)bl.sp[1].sets[bl.nset[k]].size());
        for (int i = 0; i < np; i++) rb.push_back(k);
        const int p_src = bl.sp[0].set_of_M(bl.sp[0].Ms[bl.pset[k]] - to.mu2);
        const int n_src = bl.sp[1].set_of_M(bl.sp[1].Ms[bl.nset[k]] - to.mu2);
        auto it = p_src >= 0 ? rblock.find({p_src, bl.nset[k]}) : rblock.end();
        sP.push_back(it == rblock.end() ? -1 : it->second);
        auto jt = n_src >= 0 ? rblock.find({bl.pset[k], n_src}) : rblock.end();
        sN.push_back(jt == rblock.end() ? -1 : jt->second);
    }
    for (int k = 0; k < br.nb; k++) nNr.push_back((int32_t)br.sp[1].sets[br.nset[k]].size());
    t->R = (int)rb.size();
    for (auto* x : {&rb, &r0, &nN, &ps, &ns, &sP, &sN, &nNr})
        if (x->empty()) x->push_back(0);
    t->row_block.upload(rb);
    t->block_row0.upload(r0);
    t->nN_l.upload(nN);
    t->tdims_l.upload(bl.tdims);
    t->pset_l.upload(ps);
    t->nset_l.upload(ns);
    t->srcP.upload(sP);
    t->srcN.upload(sN);
    t->nN_r.upload(nNr);
    t->tdims_r.upload(br.tdims);
    for (int op = 0; op < 2; op++)
        for (int sp = 0; sp < 2; sp++)
            upload_csr_sets(to.op[op][sp], t->rp_base[op][sp], t->ent_base[op][sp], t->rowptr[op][sp], t->col[op][sp],
                            t->val[op][sp]);
    SM_HIP(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
    return t;
}

void transitions_apply(const DeviceTransitions& t, int which, const double* v, double* w, hipStream_t s) {
    if (which != 0 && which != 1) throw std::invalid_argument("operator: 0 (M1) or 1 (E2)");
    if (t.R == 0) return;
    TrView A;
    A.row_block = t.row_block.as<int32_t>();
    A.block_row0 = t.block_row0.as<int32_t>();
    A.nN = t.nN_l.as<int32_t>();
    A.tdims = t.tdims_l.as<int64_t>();
    A.pset = t.pset_l.as<int32_t>();
    A.nset = t.nset_l.as<int32_t>();
    A.srcP = t.srcP.as<int32_t>();
    A.srcN = t.srcN.as<int32_t>();
    A.nN_r = t.nN_r.as<int32_t>();
    A.tdims_r = t.tdims_r.as<int64_t>();
    for (int sp = 0; sp < 2; sp++) {
        A.rp_base[sp] = t.rp_base[which][sp].as<int64_t>();
        A.ent_base[sp] = t.ent_base[which][sp].as<int64_t>();
        A.rowptr[sp] = t.rowptr[which][sp].as<int64_t>();
        A.col[sp] = t.col[which][sp].as<int32_t>();
        A.val[sp] = t.val[which][sp].as<double>();
    }
    hipLaunchKernelGGL(k_transition, dim3(t.R), dim3(MTB), 0, s, A, v, w);
    SM_HIP(hipGetLastError());
}

static double dot_on(DBuf& part, hipStream_t st, const double* x, const double* y, int64_t n) {
    const int G = (int)std::max<int64_t>(1, std::min<int64_t>(1024, (n + MTB - 1) / MTB));
    part.ensure((size_t)G * 8);
    hipLaunchKernelGGL(k_dot_part, dim3(G), dim3(MTB), 0, st, x, y, n, part.as<double>());
    SM_HIP(hipGetLastError());
    std::vector<double> h(G);
    SM_HIP(hipMemcpyAsync(h.data(), part.p, (size_t)G * 8, hipMemcpyDeviceToHost, st));
    SM_HIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (double x2 : h) s += x2;
    return s;
}

static bool tri_relation(int ja2, int jb2, int jc2) {
    return std::abs(ja2 - jb2) <= jc2 && jc2 <= ja2 + jb2;
}

void transit(DeviceTransitions& t, int jl2, int jr2, int nl, const double* wl, int nr, const double* wr, double* m1,
             double* e2) {
    SM_HIP(hipSetDevice(t.device));
    const double nan = std::numeric_limits<double>::quiet_NaN();
    for (int64_t k = 0; k < (int64_t)nl * nr; k++) { m1[k] = nan; e2[k] = nan; }
    if (jl2 < 0 || jr2 < 0 || nl <= 0 || nr <= 0 || t.mdim_l == 0 || t.mdim_r == 0) return;
    const int Ml2 = t.twoM_l, Mr2 = t.twoM_r;
    // the operators are components mu of rank lam (mu <= lam), and the reference's selection
    const double sel[2] = {tri_relation(jl2, jr2, 2) && t.mu2 <= 2 ? wigner3j(jl2, 2, jr2, -Ml2, t.mu2, Mr2) : 0.0,
                           tri_relation(jl2, jr2, 4) ? wigner3j(jl2, 4, jr2, -Ml2, t.mu2, Mr2) : 0.0};
    if (sel[0] == 0.0 && sel[1] == 0.0) return;
    t.left.ensure((size_t)nl * t.mdim_l * 8);
    t.vr.ensure((size_t)t.mdim_r * 8);
    t.wl.ensure((size_t)t.mdim_l * 8);
    SM_HIP(hipMemcpyAsync(t.left.p, wl, (size_t)nl * t.mdim_l * 8, hipMemcpyHostToDevice, t.stream));
    for (int r = 0; r < nr; r++) {
        SM_HIP(hipMemcpyAsync(t.vr.p, wr + (int64_t)r * t.mdim_r, (size_t)t.mdim_r * 8, hipMemcpyHostToDevice,
                              t.stream));
        for (int op = 0; op < 2; op++) {
            if (sel[op] == 0.0) continue;
            transitions_apply(t, op, t.vr.as<double>(), t.wl.as<double>(), t.stream);
            double* out = op == 0 ? m1 : e2;
            for (int l = 0; l < nl; l++)
                out[(int64_t)l * nr + r] =
                    dot_on(t.part, t.stream, t.left.as<double>() + (int64_t)l * t.mdim_l, t.wl.as<double>(), t.mdim_l) /
                    sel[op];
        }
    }
    SM_HIP(hipStreamSynchronize(t.stream));
}

}  // namespace smgpu
