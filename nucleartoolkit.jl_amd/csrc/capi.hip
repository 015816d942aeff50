// C ABI (include/smgpu.h).  Every entry point catches all exceptions and turns
// them into a status code plus a thread-local message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>

#include "../../include/smgpu.h"
#include "dense_eig.h"
#include "sm_device.h"
#include "sm_host.h"
#include "sm_status.h"
#include "sm_trl.h"
#include "tridiag.h"

using namespace smgpu;

struct smgpu_problem_s {
    Snt snt;
    std::string nucleus;
    Basis basis;
    Tables tables;
    int Z = 0, N = 0, vp = 0, vn = 0;
    int nthreads = 1;
};

struct smgpu_ham_s {
    std::unique_ptr<DeviceHam> dev;
};

struct smgpu_moments_s {
    std::unique_ptr<DeviceMoments> dev;
};

struct smgpu_transitions_s {
    std::unique_ptr<DeviceTransitions> dev;
};

template <typename TA, typename TB>
__global__ void k_cast(const TA* a, TB* b, int64_t n) {
    for (int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x; x < n; x += (int64_t)gridDim.x * 256) b[x] = (TB)a[x];
}

// two timing events, destroyed on every exit path
struct TimingEvents {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TimingEvents() {
        SM_HIP(hipEventCreate(&e0));
        try {
            SM_HIP(hipEventCreate(&e1));
        } catch (...) {
            (void)hipEventDestroy(e0);
            throw;
        }
    }
    ~TimingEvents() {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    TimingEvents(const TimingEvents&) = delete;
    TimingEvents& operator=(const TimingEvents&) = delete;
};

static thread_local std::string g_err;

void smgpu::set_last_error(const std::string& msg) { g_err = msg; }

extern "C" {

int32_t smgpu_version(void) { return 100; }
const char* smgpu_last_error(void) { return g_err.c_str(); }

int32_t smgpu_device_count(int32_t* n) {
    return guard([&]() {
        int c = 0;
        SM_HIP(hipGetDeviceCount(&c));
        *n = c;
        return SMGPU_OK;
    });
}

int32_t smgpu_problem_create(const char* snt_path, const char* nucleus, int32_t twoM, int32_t ignore_massop,
                             int32_t nthreads, smgpu_problem* out) {
    return guard([&]() -> int32_t {
        if (!snt_path || !nucleus || !out) throw std::invalid_argument("null argument");
        *out = nullptr;
        auto p = std::make_unique<smgpu_problem_s>();
        p->nucleus = nucleus;
        p->snt = read_snt(snt_path, nucleus, ignore_massop != 0);
        NucleusSpec ns = resolve_nucleus(p->snt, p->nucleus, twoM);
        p->Z = ns.Z; p->N = ns.N; p->vp = ns.vp; p->vn = ns.vn;
        int M = ns.twoM;
        p->nthreads = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
        p->basis = build_basis(p->snt, p->vp, p->vn, M);
        p->tables = build_tables(p->snt, p->basis, p->nthreads);
        *out = p.release();
        return SMGPU_OK;
    });
}

int32_t smgpu_problem_destroy(smgpu_problem p) {
    delete p;
    return SMGPU_OK;
}

int32_t smgpu_problem_set_interaction(smgpu_problem p, const char* snt_path) {
    return guard([&]() -> int32_t {
        if (!p || !snt_path) throw std::invalid_argument("null argument");
        Snt s = read_snt(snt_path, p->nucleus, false);
        if (s.lp != p->snt.lp || s.ln != p->snt.ln || s.cp != p->snt.cp || s.cn != p->snt.cn)
            throw std::invalid_argument("interaction has a different model space");
        for (int i = 0; i < s.lp; i++)
            if (s.p_sps[i].n != p->snt.p_sps[i].n || s.p_sps[i].l != p->snt.p_sps[i].l ||
                s.p_sps[i].j2 != p->snt.p_sps[i].j2)
                throw std::invalid_argument("interaction has different orbits");
        p->snt = s;
        build_coefficients(p->snt, p->basis, p->tables, p->nthreads);
        return SMGPU_OK;
    });
}

int32_t smgpu_problem_get_info(smgpu_problem p, smgpu_problem_info* info) {
    return guard([&]() -> int32_t {
        if (!p || !info) throw std::invalid_argument("null argument");
        std::memset(info, 0, sizeof(*info));
        info->Z = p->Z; info->N = p->N; info->A = p->snt.A; info->vp = p->vp; info->vn = p->vn;
        info->twoM = p->basis.twoM;
        info->lp = p->snt.lp; info->ln = p->snt.ln; info->cp = p->snt.cp; info->cn = p->snt.cn;
        info->nblocks = p->basis.nb;
        info->n_msps_p = p->basis.sp[0].L; info->n_msps_n = p->basis.sp[1].L;
        info->mdim = p->basis.mdim;
        for (int b = 0; b < p->basis.nb; b++) {
            info->n_psd += (int64_t)p->basis.sp[0].sets[p->basis.pset[b]].size();
            info->n_nsd += (int64_t)p->basis.sp[1].sets[p->basis.nset[b]].size();
        }
        info->n_pp_ops = p->tables.n_pp_ops; info->n_nn_ops = p->tables.n_nn_ops;
        info->n_pn_ops = p->tables.n_pn_ops;
        for (auto& h : p->tables.hops) info->n_hop_entries += (int64_t)h.ent.size();
        return SMGPU_OK;
    });
}

int32_t smgpu_basis_export(smgpu_problem p, int64_t* tdims, int32_t* Mp, int32_t* Mn, int64_t* pbits,
                           int64_t* psd_off, int64_t* nbits, int64_t* nsd_off) {
    return guard([&]() -> int32_t {
        if (!p) throw std::invalid_argument("null problem");
        const Basis& b = p->basis;
        int64_t po = 0, no = 0;
        for (int k = 0; k <= b.nb; k++) if (tdims) tdims[k] = b.tdims[k];
        for (int k = 0; k < b.nb; k++) {
            if (Mp) Mp[k] = b.Mp[k];
            if (Mn) Mn[k] = b.Mn[k];
            if (psd_off) psd_off[k] = po;
            if (nsd_off) nsd_off[k] = no;
            for (uint64_t x : b.sp[0].sets[b.pset[k]]) { if (pbits) pbits[po] = (int64_t)x; po++; }
            for (uint64_t x : b.sp[1].sets[b.nset[k]]) { if (nbits) nbits[no] = (int64_t)x; no++; }
        }
        if (psd_off) psd_off[b.nb] = po;
        if (nsd_off) nsd_off[b.nb] = no;
        return SMGPU_OK;
    });
}

int32_t smgpu_hops_export(smgpu_problem p, int64_t* n, int32_t* species, int32_t* bi, int32_t* bf, int32_t* i,
                          int32_t* f, int32_t* a, int32_t* c, int8_t* phase) {
    return guard([&]() -> int32_t {
        if (!p || !n) throw std::invalid_argument("null argument");
        const Basis& b = p->basis;
        const Tables& t = p->tables;
        std::vector<int> pblk(b.sp[0].sets.size(), -1), nblk(b.sp[1].sets.size(), -1);
        for (int k = 0; k < b.nb; k++) { pblk[b.pset[k]] = k; nblk[b.nset[k]] = k; }
        int64_t q = 0;
        for (const HopTable& H : t.hops) {
            int s = H.species;
            int bi_ = s == 0 ? pblk[H.iset] : nblk[H.iset];
            int bf_ = s == 0 ? pblk[H.fset] : nblk[H.fset];
            if (bi_ < 0 || bf_ < 0) continue;
            int k = t.ob[s].index_of_dM(H.dM);
            for (size_t ff = 0; ff + 1 < H.rowptr.size(); ff++)
                for (int64_t e = H.rowptr[ff]; e < H.rowptr[ff + 1]; e++) {
                    if (species) {
                        uint32_t x = H.ent[e];
                        int op = (int)((x >> 20) & 0x7FFu);
                        species[q] = s; bi[q] = bi_; bf[q] = bf_;
                        i[q] = (int32_t)(x & 0xFFFFFu); f[q] = (int32_t)ff;
                        a[q] = t.ob[s].ops[k][op].first; c[q] = t.ob[s].ops[k][op].second;
                        phase[q] = (int8_t)(((x >> 31) ^ 1u) & 1u);   // the reference's +1
                    }
                    q++;
                }
        }
        *n = q;
        return SMGPU_OK;
    });
}

int32_t smgpu_pair_export(smgpu_problem p, int32_t species, int64_t* n, int32_t* blk, int32_t* f, int32_t* i,
                          double* val) {
    return guard([&]() -> int32_t {
        if (!p || !n || species < 0 || species > 1) throw std::invalid_argument("bad argument");
        const Basis& b = p->basis;
        int64_t q = 0;
        for (int k = 0; k < b.nb; k++) {
            const PairSym& m = species == 0 ? p->tables.pp[b.pset[k]] : p->tables.nn[b.nset[k]];
            for (size_t r = 0; r + 1 < m.rowptr.size(); r++)
                for (int64_t e = m.rowptr[r]; e < m.rowptr[r + 1]; e++) {
                    if (blk) { blk[q] = k; f[q] = (int32_t)r; i[q] = m.col[e]; val[q] = m.val[e]; }
                    q++;
                }
        }
        *n = q;
        return SMGPU_OK;
    });
}

int32_t smgpu_ham_create(smgpu_problem p, int32_t device, smgpu_ham* out) {
    return guard([&]() -> int32_t {
        if (!p || !out) throw std::invalid_argument("null argument");
        *out = nullptr;
        if (p->basis.mdim == 0) throw std::invalid_argument("empty M-scheme space (mdim = 0)");
        auto h = std::make_unique<smgpu_ham_s>();
        h->dev = upload_ham(p->basis, p->tables, device);
        *out = h.release();
        return SMGPU_OK;
    });
}

int32_t smgpu_ham_create_j2(smgpu_problem p, int32_t device, smgpu_ham* out) {
    return guard([&]() -> int32_t {
        if (!p || !out) throw std::invalid_argument("null argument");
        *out = nullptr;
        if (p->basis.mdim == 0) throw std::invalid_argument("empty M-scheme space (mdim = 0)");
        Snt j2 = j2_operator(p->snt);
        Tables t = build_tables(j2, p->basis, p->nthreads);
        auto h = std::make_unique<smgpu_ham_s>();
        h->dev = upload_ham(p->basis, t, device);
        h->dev->twoJmax = p->basis.sp[0].Ms.back() + p->basis.sp[1].Ms.back();
        *out = h.release();
        return SMGPU_OK;
    });
}

int32_t smgpu_ham_destroy(smgpu_ham h) {
    return guard([&]() -> int32_t {
        if (h && h->dev) SM_HIP(hipSetDevice(h->dev->device));
        delete h;
        return SMGPU_OK;
    });
}

int32_t smgpu_ham_update_interaction(smgpu_ham h, smgpu_problem p) {
    return guard([&]() -> int32_t {
        if (!h || !p) throw std::invalid_argument("null argument");
        if (h->dev->mdim != p->basis.mdim) throw std::invalid_argument("problem does not match the handle");
        update_coefficients(*h->dev, p->basis, p->tables);
        return SMGPU_OK;
    });
}

int32_t smgpu_ham_create_variant(smgpu_ham base, smgpu_problem p, smgpu_ham* out) {
    return guard([&]() -> int32_t {
        if (!base || !p || !out) throw std::invalid_argument("null argument");
        *out = nullptr;
        if (base->dev->mdim != p->basis.mdim) throw std::invalid_argument("problem does not match the handle's basis");
        SM_HIP(hipSetDevice(base->dev->device));
        auto h = std::make_unique<smgpu_ham_s>();
        h->dev = attach_ham(base->dev->idx, p->basis, p->tables);
        *out = h.release();
        return SMGPU_OK;
    });
}

int32_t smgpu_table_bytes(smgpu_ham h, int32_t precision, int32_t max_arrays, int64_t* bytes, int32_t* n_arrays,
                          char* names_buf, int32_t names_len) {
    return guard([&]() -> int32_t {
        if (!h) throw std::invalid_argument("null handle");
        auto tb = h->dev->table_bytes(precision == SMGPU_FP64 ? 0 : 1);
        std::string names;
        for (int k = 0; k < (int)tb.size(); k++) {
            if (k < max_arrays && bytes) bytes[k] = tb[k].second;
            names += tb[k].first + "\n";
        }
        if (n_arrays) *n_arrays = (int32_t)tb.size();
        if (names_buf && names_len > 0) {
            std::strncpy(names_buf, names.c_str(), names_len - 1);
            names_buf[names_len - 1] = 0;
        }
        return SMGPU_OK;
    });
}

int32_t smgpu_lds_bytes(smgpu_ham h, int32_t precision, int64_t* parts) {
    return guard([&]() -> int32_t {
        if (!h || !parts) throw std::invalid_argument("null argument");
        const int k = precision == SMGPU_FP64 ? 0 : 1;
        const int64_t es = k == 0 ? 8 : 4;
        for (int i = 0; i < 5; i++) parts[i] = h->dev->idx->lds_elems[k][i] * es;
        return SMGPU_OK;
    });
}

int32_t smgpu_work_terms(smgpu_ham h, int64_t* t) {
    return guard([&]() -> int32_t {
        if (!h || !t) throw std::invalid_argument("null argument");
        t[0] = h->dev->pn_terms; t[1] = h->dev->pn_terms_exact; t[2] = h->dev->pp_terms;
        t[3] = h->dev->nn_terms; t[4] = h->dev->mdim;
        return SMGPU_OK;
    });
}

int32_t smgpu_apply_H(smgpu_ham h, int32_t precision, const void* v_dev, void* w_dev, void* stream) {
    return guard([&]() -> int32_t {
        if (!h || !v_dev || !w_dev) throw std::invalid_argument("null argument");
        SM_HIP(hipSetDevice(h->dev->device));
        hipStream_t s = stream ? (hipStream_t)stream : h->dev->stream;
        launch_hv(*h->dev, precision == SMGPU_FP64 ? 0 : 1, v_dev, w_dev, s);
        return SMGPU_OK;
    });
}

int32_t smgpu_apply_H_host(smgpu_ham h, int32_t precision, const double* v, double* w) {
    return guard([&]() -> int32_t {
        if (!h || !v || !w) throw std::invalid_argument("null argument");
        DeviceHam& d = *h->dev;
        SM_HIP(hipSetDevice(d.device));
        int64_t n = d.mdim;
        DBuf dv, dw;
        dv.alloc(n * 8);
        dw.alloc(n * 8);
        SM_HIP(hipMemcpyAsync(dv.p, v, n * 8, hipMemcpyHostToDevice, d.stream));
        int g = (int)std::min<int64_t>(4096, (n + 255) / 256);
        if (precision == SMGPU_FP64) {
            launch_hv(d, 0, dv.p, dw.p, d.stream);
        } else {
            DBuf fv, fw;
            fv.alloc(n * 4);
            fw.alloc(n * 4);
            hipLaunchKernelGGL((k_cast<double, float>), dim3(g), dim3(256), 0, d.stream, dv.as<double>(),
                               fv.as<float>(), n);
            launch_hv(d, 1, fv.p, fw.p, d.stream);
            hipLaunchKernelGGL((k_cast<float, double>), dim3(g), dim3(256), 0, d.stream, fw.as<float>(),
                               dw.as<double>(), n);
            SM_HIP(hipStreamSynchronize(d.stream));
        }
        SM_HIP(hipMemcpyAsync(w, dw.p, n * 8, hipMemcpyDeviceToHost, d.stream));
        SM_HIP(hipStreamSynchronize(d.stream));
        return SMGPU_OK;
    });
}

int32_t smgpu_apply_H_multi(smgpu_ham h, int32_t precision, int32_t nvec, const void* v_dev, int64_t ldv,
                            void* w_dev, int64_t ldw, void* stream) {
    return guard([&]() -> int32_t {
        if (!h || !v_dev || !w_dev || nvec < 1) throw std::invalid_argument("bad argument");
        if (nvec > 1 && (ldv < h->dev->mdim || ldw < h->dev->mdim)) throw std::invalid_argument("ld below mdim");
        SM_HIP(hipSetDevice(h->dev->device));
        hipStream_t s = stream ? (hipStream_t)stream : h->dev->stream;
        launch_hv_multi(*h->dev, precision == SMGPU_FP64 ? 0 : 1, nvec, v_dev, ldv, w_dev, ldw, s);
        return SMGPU_OK;
    });
}

int32_t smgpu_apply_H_host_multi(smgpu_ham h, int32_t precision, int32_t nvec, const double* v, double* w) {
    return guard([&]() -> int32_t {
        if (!h || !v || !w || nvec < 1) throw std::invalid_argument("bad argument");
        DeviceHam& d = *h->dev;
        SM_HIP(hipSetDevice(d.device));
        const int64_t n = d.mdim, tot = n * nvec;
        DBuf dv, dw;
        dv.alloc(tot * 8);
        dw.alloc(tot * 8);
        SM_HIP(hipMemcpyAsync(dv.p, v, tot * 8, hipMemcpyHostToDevice, d.stream));
        const int g = (int)std::min<int64_t>(4096, (tot + 255) / 256);
        if (precision == SMGPU_FP64) {
            launch_hv_multi(d, 0, nvec, dv.p, n, dw.p, n, d.stream);
        } else {
            DBuf fv, fw;
            fv.alloc(tot * 4);
            fw.alloc(tot * 4);
            hipLaunchKernelGGL((k_cast<double, float>), dim3(g), dim3(256), 0, d.stream, dv.as<double>(),
                               fv.as<float>(), tot);
            launch_hv_multi(d, 1, nvec, fv.p, n, fw.p, n, d.stream);
            hipLaunchKernelGGL((k_cast<float, double>), dim3(g), dim3(256), 0, d.stream, fw.as<float>(),
                               dw.as<double>(), tot);
            SM_HIP(hipStreamSynchronize(d.stream));
        }
        SM_HIP(hipMemcpyAsync(w, dw.p, tot * 8, hipMemcpyDeviceToHost, d.stream));
        SM_HIP(hipStreamSynchronize(d.stream));
        return SMGPU_OK;
    });
}

int32_t smgpu_time_apply_H_multi(smgpu_ham h, int32_t precision, int32_t nvec, int32_t reps, double* ms) {
    return guard([&]() -> int32_t {
        if (!h || !ms || reps < 1 || nvec < 1) throw std::invalid_argument("bad argument");
        DeviceHam& d = *h->dev;
        SM_HIP(hipSetDevice(d.device));
        const int64_t n = d.mdim, tot = n * nvec;
        const size_t s = precision == SMGPU_FP64 ? 8 : 4;
        DBuf dv, dw;
        dv.alloc(tot * s);
        dw.alloc(tot * s);
        std::vector<double> hv(tot);
        for (int64_t i = 0; i < tot; i++)
            hv[i] = 1.0 / std::sqrt((double)n) * (1.0 + 0.25 * std::sin(0.37 * (double)i));
        if (s == 8) SM_HIP(hipMemcpy(dv.p, hv.data(), tot * 8, hipMemcpyHostToDevice));
        else {
            std::vector<float> f(hv.begin(), hv.end());
            SM_HIP(hipMemcpy(dv.p, f.data(), tot * 4, hipMemcpyHostToDevice));
        }
        const int p = precision == SMGPU_FP64 ? 0 : 1;
        for (int r = 0; r < 2; r++) launch_hv_multi(d, p, nvec, dv.p, n, dw.p, n, d.stream);
        TimingEvents ev;
        hipEvent_t e0 = ev.e0, e1 = ev.e1;
        SM_HIP(hipEventRecord(e0, d.stream));
        for (int r = 0; r < reps; r++) launch_hv_multi(d, p, nvec, dv.p, n, dw.p, n, d.stream);
        SM_HIP(hipEventRecord(e1, d.stream));
        SM_HIP(hipEventSynchronize(e1));
        float t = 0.f;
        SM_HIP(hipEventElapsedTime(&t, e0, e1));
        *ms = (double)t / reps;
        return SMGPU_OK;
    });
}

// the handles' device Hamiltonians after checking that they share one index
static std::vector<const DeviceHam*> variant_hams(const smgpu_ham* hs, int32_t nvec) {
    if (!hs || nvec < 1) throw std::invalid_argument("bad argument");
    std::vector<const DeviceHam*> d(nvec);
    for (int b = 0; b < nvec; b++) {
        if (!hs[b]) throw std::invalid_argument("null handle");
        d[b] = hs[b]->dev.get();
        if (d[b]->idx != d[0]->idx) throw std::invalid_argument("handles do not share one index");
    }
    return d;
}

int32_t smgpu_apply_H_variants(const smgpu_ham* hs, int32_t precision, int32_t nvec, const void* const* v_dev,
                               void* const* w_dev, void* stream) {
    return guard([&]() -> int32_t {
        auto d = variant_hams(hs, nvec);
        if (!v_dev || !w_dev) throw std::invalid_argument("null argument");
        SM_HIP(hipSetDevice(d[0]->device));
        hipStream_t s = stream ? (hipStream_t)stream : d[0]->stream;
        launch_hv_variants(d.data(), precision == SMGPU_FP64 ? 0 : 1, nvec, v_dev, w_dev, s);
        return SMGPU_OK;
    });
}

int32_t smgpu_apply_H_host_variants(const smgpu_ham* hs, int32_t precision, int32_t nvec, const double* v,
                                    double* w) {
    return guard([&]() -> int32_t {
        auto d = variant_hams(hs, nvec);
        if (!v || !w) throw std::invalid_argument("null argument");
        const DeviceHam& h0 = *d[0];
        SM_HIP(hipSetDevice(h0.device));
        hipStream_t s = h0.stream;
        const int64_t n = h0.mdim, tot = n * nvec;
        const bool fp64 = precision == SMGPU_FP64;
        const size_t es = fp64 ? 8 : 4;
        DBuf dv, dw, fv, fw;
        dv.alloc(tot * 8);
        dw.alloc(tot * 8);
        SM_HIP(hipMemcpyAsync(dv.p, v, tot * 8, hipMemcpyHostToDevice, s));
        const int g = (int)std::min<int64_t>(4096, (tot + 255) / 256);
        if (!fp64) {
            fv.alloc(tot * 4);
            fw.alloc(tot * 4);
            hipLaunchKernelGGL((k_cast<double, float>), dim3(g), dim3(256), 0, s, dv.as<double>(), fv.as<float>(), tot);
        }
        char* vin = static_cast<char*>(fp64 ? dv.p : fv.p);
        char* wout = static_cast<char*>(fp64 ? dw.p : fw.p);
        std::vector<const void*> vp(nvec);
        std::vector<void*> wp(nvec);
        for (int b = 0; b < nvec; b++) {
            vp[b] = vin + (size_t)b * n * es;
            wp[b] = wout + (size_t)b * n * es;
        }
        launch_hv_variants(d.data(), fp64 ? 0 : 1, nvec, vp.data(), wp.data(), s);
        if (!fp64)
            hipLaunchKernelGGL((k_cast<float, double>), dim3(g), dim3(256), 0, s, fw.as<float>(), dw.as<double>(), tot);
        SM_HIP(hipMemcpyAsync(w, dw.p, tot * 8, hipMemcpyDeviceToHost, s));
        SM_HIP(hipStreamSynchronize(s));
        return SMGPU_OK;
    });
}

int32_t smgpu_time_apply_H_variants(const smgpu_ham* hs, int32_t precision, int32_t nvec, int32_t reps, double* ms) {
    return guard([&]() -> int32_t {
        auto d = variant_hams(hs, nvec);
        if (!ms || reps < 1) throw std::invalid_argument("bad argument");
        const DeviceHam& h0 = *d[0];
        SM_HIP(hipSetDevice(h0.device));
        hipStream_t st = h0.stream;
        const int64_t n = h0.mdim, tot = n * nvec;
        const bool fp64 = precision == SMGPU_FP64;
        const size_t es = fp64 ? 8 : 4;
        DBuf dv, dw;
        dv.alloc(tot * es);
        dw.alloc(tot * es);
        std::vector<double> hv(tot);
        for (int64_t i = 0; i < tot; i++)
            hv[i] = 1.0 / std::sqrt((double)n) * (1.0 + 0.25 * std::sin(0.37 * (double)i));
        if (fp64) SM_HIP(hipMemcpy(dv.p, hv.data(), tot * 8, hipMemcpyHostToDevice));
        else {
            std::vector<float> f(hv.begin(), hv.end());
            SM_HIP(hipMemcpy(dv.p, f.data(), tot * 4, hipMemcpyHostToDevice));
        }
        std::vector<const void*> vp(nvec);
        std::vector<void*> wp(nvec);
        for (int b = 0; b < nvec; b++) {
            vp[b] = static_cast<const char*>(dv.p) + (size_t)b * n * es;
            wp[b] = static_cast<char*>(dw.p) + (size_t)b * n * es;
        }
        const int p = fp64 ? 0 : 1;
        for (int r = 0; r < 2; r++) launch_hv_variants(d.data(), p, nvec, vp.data(), wp.data(), st);
        TimingEvents ev;
        hipEvent_t e0 = ev.e0, e1 = ev.e1;
        SM_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < reps; r++) launch_hv_variants(d.data(), p, nvec, vp.data(), wp.data(), st);
        SM_HIP(hipEventRecord(e1, st));
        SM_HIP(hipEventSynchronize(e1));
        float t = 0.f;
        SM_HIP(hipEventElapsedTime(&t, e0, e1));
        *ms = (double)t / reps;
        return SMGPU_OK;
    });
}

int32_t smgpu_time_apply_H(smgpu_ham h, int32_t precision, int32_t reps, double* ms) {
    return guard([&]() -> int32_t {
        if (!h || !ms || reps < 1) throw std::invalid_argument("bad argument");
        DeviceHam& d = *h->dev;
        SM_HIP(hipSetDevice(d.device));
        int64_t n = d.mdim;
        size_t s = precision == SMGPU_FP64 ? 8 : 4;
        DBuf dv, dw;
        dv.alloc(n * s);
        dw.alloc(n * s);
        std::vector<double> hv(n);
        for (int64_t i = 0; i < n; i++) hv[i] = 1.0 / std::sqrt((double)n) * (1.0 + 0.25 * std::sin(0.37 * (double)i));
        if (s == 8) SM_HIP(hipMemcpy(dv.p, hv.data(), n * 8, hipMemcpyHostToDevice));
        else {
            std::vector<float> f(hv.begin(), hv.end());
            SM_HIP(hipMemcpy(dv.p, f.data(), n * 4, hipMemcpyHostToDevice));
        }
        int p = precision == SMGPU_FP64 ? 0 : 1;
        for (int r = 0; r < 2; r++) launch_hv(d, p, dv.p, dw.p, d.stream);
        TimingEvents ev;
        hipEvent_t e0 = ev.e0, e1 = ev.e1;
        SM_HIP(hipEventRecord(e0, d.stream));
        for (int r = 0; r < reps; r++) launch_hv(d, p, dv.p, dw.p, d.stream);
        SM_HIP(hipEventRecord(e1, d.stream));
        SM_HIP(hipEventSynchronize(e1));
        float t = 0.f;
        SM_HIP(hipEventElapsedTime(&t, e0, e1));
        *ms = (double)t / reps;
        return SMGPU_OK;
    });
}

void smgpu_default_params(smgpu_solve_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->num_ev = 5; p->lm = 300; p->ls = 20; p->num_history = 3; p->tol = 1e-8;
    p->precision = SMGPU_FP64; p->refine = 1; p->max_iter = 100000; p->seed = 1234;
}

int32_t smgpu_trl_solve(smgpu_ham h, const smgpu_solve_params* params, double* evals, double* evecs,
                        smgpu_solve_stats* stats) {
    return guard([&]() -> int32_t {
        if (!h || !params || !evals) throw std::invalid_argument("null argument");
        smgpu_solve_params p = *params;
        if (p.max_iter <= 0) p.max_iter = 100000;
        if (p.num_history < 2) p.num_history = 3;
        if (p.num_ev > 64) throw std::invalid_argument("num_ev above 64");
        smgpu_solve_stats st;
        std::memset(&st, 0, sizeof(st));
        p.num_ev = (int32_t)std::min<int64_t>(p.num_ev, h->dev->mdim);
        trl_solve(*h->dev, p, evals, evecs, st);
        if (stats) *stats = st;
        return SMGPU_OK;
    });
}

int32_t smgpu_trl_solve_v0(smgpu_ham h, const smgpu_solve_params* params, const double* v0, double* evals,
                           double* evecs, smgpu_solve_stats* stats) {
    return guard([&]() -> int32_t {
        if (!h || !params || !evals) throw std::invalid_argument("null argument");
        smgpu_solve_params p = *params;
        if (p.max_iter <= 0) p.max_iter = 100000;
        if (p.num_history < 2) p.num_history = 3;
        if (p.num_ev > 64) throw std::invalid_argument("num_ev above 64");
        smgpu_solve_stats st;
        std::memset(&st, 0, sizeof(st));
        p.num_ev = (int32_t)std::min<int64_t>(p.num_ev, h->dev->mdim);
        trl_solve(*h->dev, p, evals, evecs, st, v0);
        if (stats) *stats = st;
        return SMGPU_OK;
    });
}

int32_t smgpu_trl_solve_j(smgpu_ham h, smgpu_ham j2, int32_t twoJ, const smgpu_solve_params* params,
                          const double* v0, double* evals, double* evecs, smgpu_solve_stats* stats) {
    return guard([&]() -> int32_t {
        if (!h || !j2 || !params || !evals) throw std::invalid_argument("null argument");
        if (j2->dev->twoJmax < 0) throw std::invalid_argument("j2 is not a J^2 handle (smgpu_ham_create_j2)");
        if (twoJ < 0 || twoJ > j2->dev->twoJmax) throw std::invalid_argument("target 2J outside the space");
        smgpu_solve_params p = *params;
        if (p.max_iter <= 0) p.max_iter = 100000;
        if (p.num_history < 2) p.num_history = 3;
        if (p.num_ev > 64) throw std::invalid_argument("num_ev above 64");
        smgpu_solve_stats st;
        std::memset(&st, 0, sizeof(st));
        p.num_ev = (int32_t)std::min<int64_t>(p.num_ev, h->dev->mdim);
        JProj jp;
        jp.j2 = j2->dev.get();
        jp.twoJ = twoJ;
        jp.twoJmax = j2->dev->twoJmax;
        trl_solve(*h->dev, p, evals, evecs, st, v0, nullptr, &jp);
        if (stats) *stats = st;
        return SMGPU_OK;
    });
}

// lockstep batch on d; hams: per-member Hamiltonians on d's index, or null
static int32_t solve_batch(DeviceHam& d, DeviceHam* const* hams, int32_t nb, const smgpu_solve_params* params,
                           const double* const* v0, double* const* evals, double* const* evecs,
                           smgpu_solve_stats* stats, int32_t* status) {
    SM_HIP(hipSetDevice(d.device));
    std::vector<smgpu_solve_params> p(params, params + nb);
    for (auto& q : p) {
        if (q.max_iter <= 0) q.max_iter = 100000;
        if (q.num_history < 2) q.num_history = 3;
        if (q.num_ev > 64) throw std::invalid_argument("num_ev above 64");
        q.num_ev = (int32_t)std::min<int64_t>(q.num_ev, d.mdim);
    }
    std::vector<SolveCtx> ctx(nb);
    std::vector<const SolveCtx*> cp(nb);
    struct Streams {
        std::vector<SolveCtx>& c;
        ~Streams() {
            for (auto& x : c) if (x.stream) { (void)hipStreamSynchronize(x.stream); (void)hipStreamDestroy(x.stream); }
        }
    } guard_streams{ctx};
    for (int b = 0; b < nb; b++) {
        SM_HIP(hipStreamCreateWithFlags(&ctx[b].stream, hipStreamNonBlocking));
        cp[b] = &ctx[b];
    }
    std::vector<smgpu_solve_stats> st(nb);
    std::memset(st.data(), 0, sizeof(smgpu_solve_stats) * nb);
    std::vector<int32_t> stat(nb, SMGPU_OK);
    std::vector<std::string> msgs(nb);
    BatchIO io;
    io.hams = hams;
    io.p = p.data();
    io.v0 = v0;
    io.evals = evals;
    io.evecs = evecs;
    io.st = st.data();
    io.status = stat.data();
    io.msgs = msgs.data();
    trl_solve_batch(d, nb, nb, io, cp.data(), d.stream);
    if (stats) for (int b = 0; b < nb; b++) stats[b] = st[b];
    if (status) for (int b = 0; b < nb; b++) status[b] = stat[b];
    for (int b = 0; b < nb; b++)
        if (stat[b] != SMGPU_OK) {
            set_last_error("batch member " + std::to_string(b) + ": " + msgs[b]);
            return stat[b];
        }
    return SMGPU_OK;
}

int32_t smgpu_trl_solve_batch(smgpu_ham h, int32_t nb, const smgpu_solve_params* params, const double* const* v0,
                              double* const* evals, double* const* evecs, smgpu_solve_stats* stats,
                              int32_t* status) {
    return guard([&]() -> int32_t {
        if (!h || nb < 1 || !params || !evals) throw std::invalid_argument("null argument");
        return solve_batch(*h->dev, nullptr, nb, params, v0, evals, evecs, stats, status);
    });
}

int32_t smgpu_trl_solve_variants(const smgpu_ham* hs, int32_t nb, const smgpu_solve_params* params,
                                 const double* const* v0, double* const* evals, double* const* evecs,
                                 smgpu_solve_stats* stats, int32_t* status) {
    return guard([&]() -> int32_t {
        if (nb < 1 || !params || !evals) throw std::invalid_argument("null argument");
        variant_hams(hs, nb);
        std::vector<DeviceHam*> d(nb);
        for (int b = 0; b < nb; b++) d[b] = hs[b]->dev.get();
        for (int b = 0; b < nb; b++)
            if (d[b]->mdim != d[0]->mdim) throw std::invalid_argument("handles do not share one basis");
        return solve_batch(*d[0], d.data(), nb, params, v0, evals, evecs, stats, status);
    });
}

int32_t smgpu_trbl_solve(smgpu_ham h, smgpu_ham j2, int32_t twoJ, int32_t q, const smgpu_solve_params* params,
                         const double* v0, double* evals, double* evecs, smgpu_solve_stats* stats) {
    return guard([&]() -> int32_t {
        if (!h || !params || !evals) throw std::invalid_argument("null argument");
        smgpu_solve_params p = *params;
        if (p.max_iter <= 0) p.max_iter = 100000;
        if (p.num_history < 2) p.num_history = 3;
        if (p.num_ev > 32) throw std::invalid_argument("num_ev above 32");
        smgpu_solve_stats st;
        std::memset(&st, 0, sizeof(st));
        p.num_ev = (int32_t)std::min<int64_t>(p.num_ev, h->dev->mdim);
        JProj jp;
        if (j2) {
            if (j2->dev->twoJmax < 0) throw std::invalid_argument("j2 is not a J^2 handle (smgpu_ham_create_j2)");
            if (twoJ < 0 || twoJ > j2->dev->twoJmax) throw std::invalid_argument("target 2J outside the space");
            jp.j2 = j2->dev.get();
            jp.twoJ = twoJ;
            jp.twoJmax = j2->dev->twoJmax;
        }
        trbl_solve(*h->dev, q, p, evals, evecs, st, v0, j2 ? &jp : nullptr);
        if (stats) *stats = st;
        return SMGPU_OK;
    });
}

int32_t smgpu_lanczos_steps(smgpu_ham h, int32_t precision, int32_t nsteps, const double* v0, double* alpha,
                            double* beta) {
    return guard([&]() -> int32_t {
        if (!h || !alpha || !beta) throw std::invalid_argument("null argument");
        smgpu_solve_params p;
        smgpu_default_params(&p);
        p.precision = precision;
        p.lm = nsteps + 2;
        p.ls = 1;
        p.num_ev = 1;
        lanczos_steps(*h->dev, p, nsteps, v0, alpha, beta);
        return SMGPU_OK;
    });
}

int32_t smgpu_lanczos_steps_noalloc(smgpu_ham h, int32_t precision, int32_t nsteps, const double* v0,
                                    double* alpha, double* beta) {
    return guard([&]() -> int32_t {
        if (!h || !alpha || !beta) throw std::invalid_argument("null argument");
        if (nsteps < 1) throw std::invalid_argument("need nsteps >= 1");
        smgpu_solve_params p;
        smgpu_default_params(&p);
        p.precision = precision;
        p.lm = nsteps + 2;
        p.ls = 1;
        p.num_ev = 1;
        lanczos_steps_noalloc(*h->dev, p, nsteps, v0, alpha, beta);
        return SMGPU_OK;
    });
}

int64_t smgpu_lanczos_scratch_bytes(void) { return lanczos_scratch_max(); }

int32_t smgpu_moments_create(smgpu_problem p, int32_t device, const double* gfactors, const double* effcharge,
                             double bpar, smgpu_moments* out) {
    return guard([&]() -> int32_t {
        if (!p || !out) throw std::invalid_argument("null argument");
        *out = nullptr;
        const double g0[4] = {1.0, 0.0, 5.586, -3.826}, e0[2] = {1.5, 0.5};
        const double* g = gfactors ? gfactors : g0;
        const double* e = effcharge ? effcharge : e0;
        const double b = bpar > 0.0 ? bpar : ho_bpar_main_sm(p->snt.A);
        MomentOps mo = build_moment_ops(p->snt, p->basis, b, g, e);
        auto m = std::make_unique<smgpu_moments_s>();
        m->dev = upload_moments(mo, p->basis, device);
        *out = m.release();
        return SMGPU_OK;
    });
}

int32_t smgpu_moments_destroy(smgpu_moments m) {
    return guard([&]() -> int32_t {
        if (m && m->dev) SM_HIP(hipSetDevice(m->dev->device));
        delete m;
        return SMGPU_OK;
    });
}

int32_t smgpu_moments_apply(smgpu_moments m, int32_t which, const double* v_dev, double* w_dev, void* stream) {
    return guard([&]() -> int32_t {
        if (!m || !v_dev || !w_dev) throw std::invalid_argument("null argument");
        SM_HIP(hipSetDevice(m->dev->device));
        moments_apply(*m->dev, which, v_dev, w_dev, stream ? (hipStream_t)stream : m->dev->stream);
        return SMGPU_OK;
    });
}

int32_t smgpu_moments_apply_host(smgpu_moments m, int32_t which, int32_t nvec, const double* v, double* w) {
    return guard([&]() -> int32_t {
        if (!m || !v || !w || nvec < 0) throw std::invalid_argument("bad argument");
        DeviceMoments& d = *m->dev;
        SM_HIP(hipSetDevice(d.device));
        const size_t n = (size_t)d.mdim;
        DBuf dv, dw;
        dv.alloc(std::max<size_t>(n, 1) * 8);
        dw.alloc(std::max<size_t>(n, 1) * 8);
        for (int b = 0; b < nvec; b++) {
            SM_HIP(hipMemcpy(dv.p, v + b * n, n * 8, hipMemcpyHostToDevice));
            moments_apply(d, which, dv.as<double>(), dw.as<double>(), d.stream);
            SM_HIP(hipStreamSynchronize(d.stream));
            SM_HIP(hipMemcpy(w + b * n, dw.p, n * 8, hipMemcpyDeviceToHost));
        }
        return SMGPU_OK;
    });
}

int32_t smgpu_eval_moment(smgpu_moments m, int32_t nvec, const double* vecs, const double* totJs, double* mu,
                          double* Q) {
    return guard([&]() -> int32_t {
        if (!m || !vecs || !totJs || !mu || !Q || nvec < 0) throw std::invalid_argument("bad argument");
        eval_moment(*m->dev, nvec, vecs, totJs, mu, Q);
        return SMGPU_OK;
    });
}

int32_t smgpu_moments_bpar(smgpu_moments m, double* bpar) {
    return guard([&]() -> int32_t {
        if (!m || !bpar) throw std::invalid_argument("null argument");
        *bpar = m->dev->bpar;
        return SMGPU_OK;
    });
}

int32_t smgpu_transitions_create(smgpu_problem left, smgpu_problem right, int32_t device, const double* gfactors,
                                 const double* effcharge, double bpar, smgpu_transitions* out) {
    return guard([&]() -> int32_t {
        if (!left || !right || !out) throw std::invalid_argument("null argument");
        *out = nullptr;
        auto same_orbits = [](const std::vector<Orbit>& a, const std::vector<Orbit>& b) {
            if (a.size() != b.size()) return false;
            for (size_t i = 0; i < a.size(); i++)
                if (a[i].n != b[i].n || a[i].l != b[i].l || a[i].j2 != b[i].j2 || a[i].tz != b[i].tz) return false;
            return true;
        };
        // the right basis' m-state bits must name the same orbits as the left's
        if (left->Z != right->Z || left->N != right->N || left->snt.A != right->snt.A ||
            !same_orbits(left->snt.p_sps, right->snt.p_sps) || !same_orbits(left->snt.n_sps, right->snt.n_sps))
            throw std::invalid_argument("left and right problems of different nuclei or spaces");
        const double g0[4] = {1.0, 0.0, 5.586, -3.826}, e0[2] = {1.5, 0.5};
        const double* g = gfactors ? gfactors : g0;
        const double* e = effcharge ? effcharge : e0;
        const double b = bpar > 0.0 ? bpar : ho_bpar_main_sm(left->snt.A);
        TransitionOps to = build_transition_ops(left->snt, left->basis, right->basis, b, g, e);
        auto t = std::make_unique<smgpu_transitions_s>();
        t->dev = upload_transitions(to, left->basis, right->basis, device);
        *out = t.release();
        return SMGPU_OK;
    });
}

int32_t smgpu_transitions_destroy(smgpu_transitions t) {
    return guard([&]() -> int32_t {
        if (t && t->dev) SM_HIP(hipSetDevice(t->dev->device));
        delete t;
        return SMGPU_OK;
    });
}

int32_t smgpu_transitions_apply(smgpu_transitions t, int32_t which, const double* v_dev, double* w_dev, void* stream) {
    return guard([&]() -> int32_t {
        if (!t || !v_dev || !w_dev) throw std::invalid_argument("null argument");
        SM_HIP(hipSetDevice(t->dev->device));
        transitions_apply(*t->dev, which, v_dev, w_dev, stream ? (hipStream_t)stream : t->dev->stream);
        return SMGPU_OK;
    });
}

int32_t smgpu_transitions_apply_host(smgpu_transitions t, int32_t which, int32_t nvec, const double* v, double* w) {
    return guard([&]() -> int32_t {
        if (!t || !v || !w || nvec < 0) throw std::invalid_argument("bad argument");
        DeviceTransitions& d = *t->dev;
        SM_HIP(hipSetDevice(d.device));
        const size_t nr = (size_t)d.mdim_r, nl = (size_t)d.mdim_l;
        DBuf dv, dw;
        dv.alloc(std::max<size_t>(nr, 1) * 8);
        dw.alloc(std::max<size_t>(nl, 1) * 8);
        for (int b = 0; b < nvec; b++) {
            SM_HIP(hipMemcpy(dv.p, v + b * nr, nr * 8, hipMemcpyHostToDevice));
            transitions_apply(d, which, dv.as<double>(), dw.as<double>(), d.stream);
            SM_HIP(hipStreamSynchronize(d.stream));
            SM_HIP(hipMemcpy(w + b * nl, dw.p, nl * 8, hipMemcpyDeviceToHost));
        }
        return SMGPU_OK;
    });
}

int32_t smgpu_transit(smgpu_transitions t, int32_t jl2, int32_t jr2, int32_t nl, const double* wl, int32_t nr,
                      const double* wr, double* m1, double* e2) {
    return guard([&]() -> int32_t {
        if (!t || !m1 || !e2 || nl < 0 || nr < 0 || (nl > 0 && !wl) || (nr > 0 && !wr))
            throw std::invalid_argument("bad argument");
        transit(*t->dev, jl2, jr2, nl, wl, nr, wr, m1, e2);
        return SMGPU_OK;
    });
}

int32_t smgpu_transitions_info(smgpu_transitions t, int32_t* mu2, double* bpar, int64_t* mdim_l, int64_t* mdim_r) {
    return guard([&]() -> int32_t {
        if (!t) throw std::invalid_argument("null argument");
        if (mu2) *mu2 = t->dev->mu2;
        if (bpar) *bpar = t->dev->bpar;
        if (mdim_l) *mdim_l = t->dev->mdim_l;
        if (mdim_r) *mdim_r = t->dev->mdim_r;
        return SMGPU_OK;
    });
}

int32_t smgpu_malloc(smgpu_ham h, int64_t bytes, void** dev) {
    return guard([&]() -> int32_t {
        if (!h || !dev || bytes < 0) throw std::invalid_argument("bad argument");
        SM_HIP(hipSetDevice(h->dev->device));
        SM_HIP(hipMalloc(dev, (size_t)std::max<int64_t>(bytes, 1)));
        return SMGPU_OK;
    });
}

int32_t smgpu_free(smgpu_ham h, void* dev) {
    return guard([&]() -> int32_t {
        if (!h) throw std::invalid_argument("null handle");
        SM_HIP(hipSetDevice(h->dev->device));
        if (dev) SM_HIP(hipFree(dev));
        return SMGPU_OK;
    });
}

int32_t smgpu_memcpy(smgpu_ham h, void* dst, const void* src, int64_t bytes, int32_t kind) {
    return guard([&]() -> int32_t {
        if (!h || !dst || !src || bytes < 0) throw std::invalid_argument("bad argument");
        hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
        SM_HIP(hipSetDevice(h->dev->device));
        SM_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, k, h->dev->stream));
        SM_HIP(hipStreamSynchronize(h->dev->stream));
        return SMGPU_OK;
    });
}

int32_t smgpu_synchronize(smgpu_ham h) {
    return guard([&]() -> int32_t {
        if (!h) throw std::invalid_argument("null handle");
        SM_HIP(hipStreamSynchronize(h->dev->stream));
        return SMGPU_OK;
    });
}

int32_t smgpu_selftest_tridiag(const double* alpha, const double* beta, int32_t n, int32_t k, double* vals,
                               double* vecs) {
    return guard([&]() -> int32_t {
        if (!alpha || !vals || !vecs || n < 1 || k < 1 || k > n) throw std::invalid_argument("bad argument");
        std::vector<double> work(5 * (size_t)n + 8);
        std::vector<double> b(std::max(1, n));
        for (int i = 0; i < n - 1; i++) b[i] = beta[i];
        tri_lowest(alpha, b.data(), n, k, vals, vecs, work.data());
        return SMGPU_OK;
    });
}

int32_t smgpu_selftest_symeig(const double* A, int32_t n, int32_t k, double* vals, double* vecs) {
    return guard([&]() -> int32_t {
        if (!A || !vals || n < 1 || k < 1 || k > n) throw std::invalid_argument("bad argument");
        DenseEig e;
        if (vecs) {
            e.reduce(A, n, n);
            e.lowest_pairs(k, vals, vecs);
        } else {
            e.lowest_values(A, n, n, k, vals);     // the band reduction when A is banded
        }
        return SMGPU_OK;
    });
}

int32_t smgpu_selftest_restart(int32_t device, const double* alpha, const double* beta, int32_t m, int32_t ls,
                               double* S, double* alpha_out, double* beta_out) {
    return guard([&]() -> int32_t {
        if (!alpha || !beta || !S || !alpha_out || !beta_out || m < 2 || ls < 1 || ls >= m || ls > 32)
            throw std::invalid_argument("bad argument");
        if (device < 0) {
            std::vector<double> a(alpha, alpha + m), b(beta, beta + m);
            std::vector<double> work(5 * (size_t)m + (size_t)ls * (m + 3 * ls + 8) + 64);
            tri_restart(a.data(), b.data(), m, ls, S, work.data());
            for (int i = 0; i < ls; i++) { alpha_out[i] = a[i]; beta_out[i] = b[i]; }
        } else {
            restart_small_device(device, alpha, beta, m, ls, S, alpha_out, beta_out);
        }
        return SMGPU_OK;
    });
}

}  // extern "C"
