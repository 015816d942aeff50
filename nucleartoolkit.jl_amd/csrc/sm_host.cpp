// Host-side construction of the basis and of the device-layout Hamiltonian
// tables.  Semantics follow NuclearToolkit.jl (paths relative to its root):
//   readsmsnt              src/ShellModel/shellmodel_main.jl:293-365
//   construct_msps / occ   shellmodel_main.jl:373-418, 732-848, 899-907
//   prep_Hamil_T1          shellmodel_main.jl:580-649
//   prep_Hamil_pn          shellmodel_main.jl:658-730
//   hop validity / phases  shellmodel_main.jl:850-897
// The layout (gather-form hop tables, merged symmetric pp/nn rows, V_eff
// matrices per dM) is this library's own; see sm_host.h.
#include "sm_host.h"

#include <quadmath.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <fstream>
#include <functional>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <thread>
#include <unordered_map>

namespace smgpu {

static void parallel_for(int64_t n, int nthreads, const std::function<void(int64_t)>& fn) {
    if (nthreads <= 1 || n <= 1) {
        for (int64_t i = 0; i < n; i++) fn(i);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> th;
    int nt = (int)std::min<int64_t>(nthreads, n);
    for (int t = 0; t < nt; t++)
        th.emplace_back([&]() {
            for (;;) {
                int64_t i = next.fetch_add(1);
                if (i >= n) break;
                fn(i);
            }
        });
    for (auto& x : th) x.join();
}

// ------------------------------------------------------------------ parsing
static std::vector<std::string> split_ws(const std::string& s) {
    std::istringstream is(s);
    std::vector<std::string> out;
    std::string t;
    while (is >> t) out.push_back(t);
    return out;
}

static const char* kElements[] = {
    "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar",
    "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br",
    "Kr", "Rb", "Sr", "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te",
    "I", "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm",
    "Yb", "Lu", "Hf", "Ta", "W", "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn"};

int element_z(const std::string& symbol) {
    for (int i = 0; i < (int)(sizeof(kElements) / sizeof(kElements[0])); i++)
        if (symbol == kElements[i]) return i + 1;
    return -1;
}

NucleusSpec resolve_nucleus(const Snt& snt, const std::string& nucleus, int twoM) {
    std::string el;
    for (char c : nucleus) if (!isdigit((unsigned char)c)) el += c;
    NucleusSpec r{};
    r.Z = element_z(el);
    if (r.Z < 0) throw std::invalid_argument("unknown element in nucleus '" + nucleus + "'");
    r.N = snt.A - r.Z;
    r.vp = r.Z - snt.cp;
    r.vn = r.N - snt.cn;
    if (r.vp < 0 || r.vn < 0) throw std::invalid_argument("nucleus below the core of the interaction");
    r.twoM = twoM < 0 ? snt.A % 2 : twoM;
    if ((r.twoM - snt.A) % 2 != 0) throw std::invalid_argument("2M parity does not match the mass number");
    return r;
}

Snt read_snt(const std::string& path, const std::string& nucleus, bool ignore_massop) {
    Snt s;
    // mass number = first run of digits (util_func.jl:1 `reg`)
    std::string digits;
    for (char c : nucleus) {
        if (isdigit((unsigned char)c)) digits += c;
        else if (!digits.empty()) break;
    }
    if (digits.empty()) throw std::runtime_error("nucleus '" + nucleus + "' has no mass number");
    s.A = std::stoi(digits);
    if (s.A <= 0) throw std::runtime_error("invalid nucleus " + nucleus);
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open interaction file " + path);
    std::vector<std::string> lines;
    std::string line;
    while (std::getline(f, line)) {
        size_t b = line.find_first_not_of(" \t\r\n");
        std::string t = b == std::string::npos ? std::string() : line.substr(b);
        if (!t.empty() && (t[0] == '!' || t[0] == '#')) continue;
        lines.push_back(t);
    }
    auto need = [&](size_t i) -> std::vector<std::string> {
        if (i >= lines.size()) throw std::runtime_error("truncated interaction file " + path);
        return split_ws(lines[i]);
    };
    try {
        auto h = need(0);
        if (h.size() < 4) throw std::runtime_error("bad model-space header");
        s.lp = std::stoi(h[0]); s.ln = std::stoi(h[1]); s.cp = std::stoi(h[2]); s.cn = std::stoi(h[3]);
        if (s.lp <= 0 || s.ln <= 0 || s.lp > 16 || s.ln > 16) throw std::runtime_error("bad orbit counts");
        s.p_sps.resize(s.lp); s.n_sps.resize(s.ln);
        for (int i = 0; i < s.lp; i++) {
            auto t = need(1 + i);
            if (t.size() < 5) throw std::runtime_error("bad orbit line");
            int ith = std::stoi(t[0]);
            if (ith < 1 || ith > s.lp) throw std::runtime_error("bad proton orbit index");
            s.p_sps[ith - 1] = {std::stoi(t[1]), std::stoi(t[2]), std::stoi(t[3]), std::stoi(t[4])};
        }
        for (int i = 0; i < s.ln; i++) {
            auto t = need(1 + i + s.ln);
            if (t.size() < 5) throw std::runtime_error("bad orbit line");
            int ith = std::stoi(t[0]) - s.lp;
            if (ith < 1 || ith > s.ln) throw std::runtime_error("bad neutron orbit index");
            s.n_sps[ith - 1] = {std::stoi(t[1]), std::stoi(t[2]), std::stoi(t[3]), std::stoi(t[4])};
        }
        size_t o = 1 + s.ln + s.lp;
        auto t = need(o);
        int nsp = std::stoi(t[0]);
        s.spe_p.assign(s.lp, 0.0); s.spe_n.assign(s.ln, 0.0);
        for (int i = 0; i < nsp; i++) {
            auto u = need(o + 1 + i);
            if (u.size() < 3) throw std::runtime_error("bad SPE line");
            int j = std::stoi(u[0]);
            if (j < 1 || j > s.lp + s.ln) throw std::runtime_error("bad SPE orbit");
            if (j <= s.lp) s.spe_p[j - 1] = std::stod(u[2]);
            else s.spe_n[j - s.lp - 1] = std::stod(u[2]);
        }
        auto u = need(o + 1 + nsp);
        if (u.size() < 3) throw std::runtime_error("TBME header needs at least 3 fields");
        int ntbme = std::stoi(u[0]);
        s.massop = std::stoi(u[1]);
        s.Aref = std::stod(u[2]);
        s.pw = u.size() > 3 ? std::stod(u[3]) : 0.0;
        for (int ith = 1; ith <= ntbme; ith++) {
            auto r = need(o + 1 + nsp + ith);
            if (r.size() < 6) throw std::runtime_error("bad TBME line");
            int i = std::stoi(r[0]), j = std::stoi(r[1]), k = std::stoi(r[2]), l = std::stoi(r[3]);
            int J = std::stoi(r[4]);
            int ch;
            if (i <= s.lp && j <= s.lp) ch = 0;
            else if (i > s.lp && j > s.lp) ch = 1;
            else if (i <= s.lp && j > s.lp) ch = 2;
            else throw std::runtime_error("TBME line with neutron before proton index");
            double v = std::stod(r[5]);
            if (s.massop == 1 && !ignore_massop) v *= std::pow((double)s.A / s.Aref, s.pw);
            bool same = (i == j) ? (k == l && k == i) : (k != l && ((i == k && j == l)));
            // readsmsnt compares unique([i,j]) with unique([k,l]) as ordered lists
            std::vector<int> u1{i}, u2{k};
            if (j != i) u1.push_back(j);
            if (l != k) u2.push_back(l);
            same = (u1 == u2);
            if (!same) { s.labels[ch].push_back({k, l, i, j, J, ith}); s.tbmes[ch].push_back(v); }
            s.labels[ch].push_back({i, j, k, l, J, ith});
            s.tbmes[ch].push_back(v);
        }
    } catch (std::invalid_argument&) {
        throw std::runtime_error("unparsable number in " + path);
    } catch (std::out_of_range&) {
        throw std::runtime_error("number out of range in " + path);
    }
    return s;
}

Snt j2_operator(const Snt& model) {
    Snt s;
    s.lp = model.lp; s.ln = model.ln; s.cp = model.cp; s.cn = model.cn;
    s.A = model.A; s.massop = 0; s.Aref = model.A; s.pw = 0.0;
    s.p_sps = model.p_sps; s.n_sps = model.n_sps;
    auto jj = [](int j2) { return 0.5 * j2 * (0.5 * j2 + 1.0); };
    s.spe_p.resize(s.lp); s.spe_n.resize(s.ln);
    for (int a = 0; a < s.lp; a++) s.spe_p[a] = jj(s.p_sps[a].j2);
    for (int a = 0; a < s.ln; a++) s.spe_n[a] = jj(s.n_sps[a].j2);
    int line = 0;
    // like particles: a <= b, J in |ja-jb|..ja+jb, J even when a == b (antisymmetry)
    for (int ch = 0; ch < 2; ch++) {
        const auto& sps = ch == 0 ? s.p_sps : s.n_sps;
        const int off = ch == 0 ? 0 : s.lp;
        for (int a = 0; a < (int)sps.size(); a++)
            for (int b = a; b < (int)sps.size(); b++) {
                const int ja = sps[a].j2, jb = sps[b].j2;
                for (int J = std::abs(ja - jb) / 2; J <= (ja + jb) / 2; J++) {
                    if (a == b && J % 2 == 1) continue;
                    s.labels[ch].push_back({a + 1 + off, b + 1 + off, a + 1 + off, b + 1 + off, J, ++line});
                    s.tbmes[ch].push_back(J * (J + 1.0) - jj(ja) - jj(jb));
                }
            }
    }
    // proton-neutron: every (a_p, b_n, J)
    for (int a = 0; a < s.lp; a++)
        for (int b = 0; b < s.ln; b++) {
            const int ja = s.p_sps[a].j2, jb = s.n_sps[b].j2;
            for (int J = std::abs(ja - jb) / 2; J <= (ja + jb) / 2; J++) {
                s.labels[2].push_back({a + 1, b + 1 + s.lp, a + 1, b + 1 + s.lp, J, ++line});
                s.tbmes[2].push_back(J * (J + 1.0) - jj(ja) - jj(jb));
            }
        }
    return s;
}

// ------------------------------------------------------------------ CG
static __float128 qfact(int n) {
    __float128 r = 1;
    for (int i = 2; i <= n; i++) r *= i;
    return r;
}

static double clebsch_gordan_eval(int j1, int m1, int j2, int m2, int J, int M);

// Every coefficient with 2j1, 2j2 <= CG_JMAX is computed once per process (about
// 4e4 non-zero values, tens of ms); the hot loops of the table build then read it.
static constexpr int CG_JMAX = 11;
static constexpr int CG_NJ = CG_JMAX + 1, CG_NJJ = 2 * CG_JMAX + 1;
static const double* cg_table() {
    static std::vector<double> tab;
    static std::once_flag once;
    std::call_once(once, [] {
        tab.assign((size_t)CG_NJ * CG_NJ * CG_NJJ * CG_NJ * CG_NJ, 0.0);
        for (int j1 = 0; j1 <= CG_JMAX; j1++)
            for (int j2 = 0; j2 <= CG_JMAX; j2++)
                for (int J = std::abs(j1 - j2); J <= j1 + j2; J += 2)
                    for (int m1 = -j1; m1 <= j1; m1 += 2)
                        for (int m2 = -j2; m2 <= j2; m2 += 2) {
                            if (std::abs(m1 + m2) > J) continue;
                            size_t i = ((((size_t)j1 * CG_NJ + j2) * CG_NJJ + J) * CG_NJ + (m1 + j1) / 2) * CG_NJ +
                                       (m2 + j2) / 2;
                            tab[i] = clebsch_gordan_eval(j1, m1, j2, m2, J, m1 + m2);
                        }
    });
    return tab.data();
}

double clebsch_gordan(int j1, int m1, int j2, int m2, int J, int M) {
    if (m1 + m2 != M || std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(M) > J) return 0.0;
    if (J > j1 + j2 || J < std::abs(j1 - j2)) return 0.0;
    if (((j1 + m1) & 1) || ((j2 + m2) & 1) || ((J + M) & 1) || ((j1 + j2 + J) & 1)) return 0.0;
    if (j1 <= CG_JMAX && j2 <= CG_JMAX)
        return cg_table()[((((size_t)j1 * CG_NJ + j2) * CG_NJJ + J) * CG_NJ + (m1 + j1) / 2) * CG_NJ + (m2 + j2) / 2];
    return clebsch_gordan_eval(j1, m1, j2, m2, J, M);
}

// correctly rounded <j1 m1; j2 m2 | J M> (doubled arguments): Racah's formula in
// quad precision, rounded once to double
static double clebsch_gordan_eval(int j1, int m1, int j2, int m2, int J, int M) {
    if (m1 + m2 != M || std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(M) > J) return 0.0;
    if (J > j1 + j2 || J < std::abs(j1 - j2)) return 0.0;
    if (((j1 + m1) & 1) || ((j2 + m2) & 1) || ((J + M) & 1) || ((j1 + j2 + J) & 1)) return 0.0;
    int a = (j1 + j2 - J) / 2, b = (j1 - j2 + J) / 2, c = (-j1 + j2 + J) / 2, d = (j1 + j2 + J) / 2 + 1;
    __float128 pref = (__float128)(J + 1) * qfact(a) * qfact(b) * qfact(c) / qfact(d);
    pref *= qfact((j1 + m1) / 2) * qfact((j1 - m1) / 2) * qfact((j2 + m2) / 2) * qfact((j2 - m2) / 2) *
            qfact((J + M) / 2) * qfact((J - M) / 2);
    __float128 s = 0;
    for (int k = 0; k <= a; k++) {
        int e[6] = {k, a - k, (j1 - m1) / 2 - k, (j2 + m2) / 2 - k, (J - j2 + m1) / 2 + k, (J - j1 - m2) / 2 + k};
        bool ok = true;
        for (int q = 0; q < 6; q++) ok = ok && e[q] >= 0;
        if (!ok) continue;
        __float128 den = 1;
        for (int q = 0; q < 6; q++) den *= qfact(e[q]);
        s += ((k & 1) ? -1 : 1) / den;
    }
    if (s == 0) return 0.0;
    __float128 r = sqrtq(pref) * s;
    return (double)r;
}

// ------------------------------------------------------------------ basis
int Species::set_of_M(int M) const {
    auto it = std::lower_bound(Ms.begin(), Ms.end(), M);
    if (it == Ms.end() || *it != M) return -1;
    return (int)(it - Ms.begin());
}

int64_t Species::find(int set, uint64_t sd) const {
    const auto& v = sets[set];
    int64_t lo = 0, hi = (int64_t)v.size() - 1;   // descending order
    while (lo <= hi) {
        int64_t mid = (lo + hi) / 2;
        if (v[mid] > sd) lo = mid + 1;
        else if (v[mid] < sd) hi = mid - 1;
        else return mid;
    }
    return -1;
}

void build_species(Species& s, const std::vector<Orbit>& sps, int nv) {
    s.orb.clear(); s.m2.clear(); s.j2.clear();
    for (int o = 0; o < (int)sps.size(); o++)
        for (int m = -sps[o].j2; m <= sps[o].j2; m += 2) {
            s.orb.push_back(o); s.m2.push_back(m); s.j2.push_back(sps[o].j2);
        }
    s.L = (int)s.orb.size();
    s.nv = nv;
    s.norb = (int)sps.size();
    if (s.L > 63) throw std::runtime_error("more than 63 m-states per species");
    if (nv < 0 || nv > s.L) throw std::runtime_error("valence particle count out of range");
    // lexicographic combinations of m-state indices = descending bit integers
    std::map<int, std::vector<uint64_t>> byM;
    std::vector<int> comb(nv);
    for (int i = 0; i < nv; i++) comb[i] = i;
    for (;;) {
        uint64_t bits = 0;
        int M = 0;
        for (int k : comb) { bits |= 1ull << (s.L - 1 - k); M += s.m2[k]; }
        byM[M].push_back(bits);
        int i = nv - 1;
        while (i >= 0 && comb[i] == s.L - nv + i) i--;
        if (i < 0) break;
        comb[i]++;
        for (int k = i + 1; k < nv; k++) comb[k] = comb[k - 1] + 1;
    }
    s.Ms.clear(); s.sets.clear(); s.jocc.clear();
    for (auto& kv : byM) {
        s.Ms.push_back(kv.first);
        s.sets.push_back(kv.second);
        std::vector<double> jo(kv.second.size() * s.norb, 0.0);
        for (size_t i = 0; i < kv.second.size(); i++)
            for (int k = 0; k < s.L; k++)
                if (kv.second[i] >> (s.L - 1 - k) & 1ull) jo[i * s.norb + s.orb[k]] += 1.0;
        s.jocc.push_back(std::move(jo));
    }
}

Basis build_basis(const Snt& snt, int vp, int vn, int twoM) {
    Basis b;
    build_species(b.sp[0], snt.p_sps, vp);
    build_species(b.sp[1], snt.n_sps, vn);
    b.twoM = twoM;
    b.tdims.push_back(0);
    for (int i = 0; i < (int)b.sp[0].Ms.size(); i++) {     // Mp ascending (occ :802-811)
        int Mp = b.sp[0].Ms[i];
        int j = b.sp[1].set_of_M(twoM - Mp);
        if (j < 0) continue;
        b.Mp.push_back(Mp); b.Mn.push_back(twoM - Mp);
        b.pset.push_back(i); b.nset.push_back(j);
        b.tdims.push_back(b.tdims.back() + (int64_t)b.sp[0].sets[i].size() * (int64_t)b.sp[1].sets[j].size());
    }
    b.nb = (int)b.Mp.size();
    b.mdim = b.tdims.back();
    return b;
}

// ------------------------------------------------------------------ operator lists
namespace {

struct Op4 { uint64_t a, b, c, d; };
struct Op4Hash {
    size_t operator()(const std::array<uint64_t, 4>& k) const {
        uint64_t h = k[0] * 0x9E3779B97F4A7C15ull;
        h ^= k[1] + 0x632BE59BD9B4E019ull + (h << 6);
        h ^= k[2] * 0xC2B2AE3D27D4EB4Full + (h >> 3);
        h ^= k[3] * 0x165667B19E3779F9ull + (h << 7);
        return (size_t)h;
    }
};

struct OpList {
    std::vector<Op4> ops;
    std::vector<double> V;
    std::unordered_map<std::array<uint64_t, 4>, size_t, Op4Hash> seen;
    void add(uint64_t a, uint64_t b, uint64_t c, uint64_t d, double v) {
        std::array<uint64_t, 4> k{a, b, c, d};
        auto it = seen.find(k);
        if (it == seen.end()) { seen.emplace(k, ops.size()); ops.push_back({a, b, c, d}); V.push_back(v); }
        else V[it->second] += v;
    }
};

// m-state indices of orbit `o` in ascending m (possible_mz, :552-565)
std::vector<int> mstates_of(const Species& s, int o) {
    std::vector<int> r;
    for (int k = 0; k < s.L; k++) if (s.orb[k] == o) r.push_back(k);
    return r;
}

inline uint64_t bit_of(const Species& s, int k) { return 1ull << (s.L - 1 - k); }

// like-particle two-body operators of one species (prep_Hamil_T1 :580-649)
OpList t1_ops(const Snt& snt, const Species& s, int vr) {
    OpList L;
    int lo = vr == 0 ? 0 : snt.lp;
    for (size_t q = 0; q < snt.labels[vr].size(); q++) {
        const auto& lab = snt.labels[vr][q];
        double ME = snt.tbmes[vr][q];
        int a = lab[0] - lo - 1, b = lab[1] - lo - 1, c = lab[2] - lo - 1, d = lab[3] - lo - 1;
        int J2 = 2 * lab[4];
        auto A = mstates_of(s, a), B = mstates_of(s, b), C = mstates_of(s, c), D = mstates_of(s, d);
        int ja = s.j2[A[0]], jb = s.j2[B[0]], jc = s.j2[C[0]], jd = s.j2[D[0]];
        double nfac = std::sqrt((1.0 + (a == b ? 1.0 : 0.0)) * (1.0 + (c == d ? 1.0 : 0.0)));
        for (int kc : C)
            for (int kd : D) {
                int mc = s.m2[kc], md = s.m2[kd];
                if (c == d && mc >= md) continue;
                if (std::abs(mc + md) > J2) continue;
                int M = mc + md;
                for (int ka : A)
                    for (int kb : B) {
                        int ma = s.m2[ka], mb = s.m2[kb];
                        if (a == b && ma >= mb) continue;
                        if (ma + mb != M) continue;
                        double cg1 = clebsch_gordan(ja, ma, jb, mb, J2, M);
                        double cg2 = clebsch_gordan(jc, mc, jd, md, J2, M);
                        double v = ME * nfac * cg1 * cg2;
                        L.add(bit_of(s, ka), bit_of(s, kb), bit_of(s, kc), bit_of(s, kd), v);
                    }
            }
    }
    return L;
}

// pn operators per proton dM (prep_Hamil_pn :658-730); key is dM
std::map<int, OpList> pn_ops(const Snt& snt, const Species& P, const Species& N) {
    std::map<int, OpList> out;
    for (size_t q = 0; q < snt.labels[2].size(); q++) {
        const auto& lab = snt.labels[2][q];
        double ME = snt.tbmes[2][q];
        int a = lab[0] - 1, c = lab[2] - 1, b = lab[1] - snt.lp - 1, d = lab[3] - snt.lp - 1;
        int J2 = 2 * lab[4];
        auto A = mstates_of(P, a), C = mstates_of(P, c), B = mstates_of(N, b), D = mstates_of(N, d);
        int ja = P.j2[A[0]], jc = P.j2[C[0]], jb = N.j2[B[0]], jd = N.j2[D[0]];
        for (int kc : C) {
            int mc = P.m2[kc];
            for (int ka : A) {
                int ma = P.m2[ka];
                int Mp = ma - mc;
                OpList& L = out[Mp];
                for (int kd : D) {
                    int md = N.m2[kd];
                    if (std::abs(mc + md) > J2) continue;
                    double cg1 = clebsch_gordan(jc, mc, jd, md, J2, mc + md);
                    double tfac = ME * cg1;
                    for (int kb : B) {
                        int mb = N.m2[kb];
                        if (mb - md + Mp != 0) continue;
                        double fac = tfac * clebsch_gordan(ja, ma, jb, mb, J2, ma + mb);
                        L.add(bit_of(P, ka), bit_of(N, kb), bit_of(P, kc), bit_of(N, kd), fac);
                    }
                }
            }
        }
    }
    return out;
}

inline int popc(uint64_t x) { return __builtin_popcountll(x); }
inline uint64_t at_or_above(uint64_t bit) { return ~(bit - 1); }   // bits of equal or higher significance

// a+ b+ d c on sd: returns false if it vanishes; sign exponent parity in *neg
bool apply2(uint64_t sd, const Op4& o, uint64_t* out, int* neg) {
    if ((sd & o.c) == 0 || (sd & o.d) == 0) return false;
    uint64_t rest = sd & ~o.c & ~o.d;
    if ((rest & o.a) || (rest & o.b)) return false;
    int e = popc(sd & at_or_above(o.c));
    uint64_t x = sd & ~o.c;
    e += popc(x & at_or_above(o.d));
    x &= ~o.d;
    e += popc(x & at_or_above(o.b));
    x |= o.b;
    e += popc(x & at_or_above(o.a));
    x |= o.a;
    *out = x;
    *neg = e & 1;
    return true;
}

}  // namespace

int OneBodyOps::index_of_dM(int dM) const {
    auto it = std::lower_bound(dMs.begin(), dMs.end(), dM);
    if (it == dMs.end() || *it != dM) return -1;
    return (int)(it - dMs.begin());
}

static void build_onebody_ops(const Species& s, OneBodyOps& ob) {
    std::map<int, std::vector<std::pair<int, int>>> m;
    for (int a = 0; a < s.L; a++)
        for (int c = 0; c < s.L; c++) m[s.m2[a] - s.m2[c]].push_back({a, c});
    ob.dMs.clear(); ob.ops.clear();
    for (auto& kv : m) { ob.dMs.push_back(kv.first); ob.ops.push_back(kv.second); }
}

// gather-form hop table: for each f in fset, inputs i = c+ a f  (so that a+ c i = +-f)
static HopTable build_hops(const Species& s, const OneBodyOps& ob, int sp, int iset, int fset) {
    HopTable h;
    h.species = sp; h.iset = iset; h.fset = fset;
    h.dM = s.Ms[fset] - s.Ms[iset];
    int k = ob.index_of_dM(h.dM);
    const auto& F = s.sets[fset];
    h.rowptr.assign(F.size() + 1, 0);
    if (k < 0) return h;
    const auto& ops = ob.ops[k];
    for (size_t f = 0; f < F.size(); f++) {
        uint64_t sdf = F[f];
        for (size_t q = 0; q < ops.size(); q++) {
            uint64_t ba = bit_of(s, ops[q].first), bc = bit_of(s, ops[q].second);
            if ((sdf & ba) == 0) continue;
            uint64_t mid = sdf & ~ba;
            if (mid & bc) continue;
            uint64_t sdi = mid | bc;
            int64_t i = s.find(iset, sdi);
            if (i < 0) continue;
            // sign of a+ c acting on sdi (calc_phase_1! without its constant
            // extra factor, which cancels between the proton and neutron parts)
            int e = popc(sdi & at_or_above(bc));
            uint64_t x = sdi & ~bc;
            e += popc(x & at_or_above(ba));
            if ((uint64_t)i >= (1u << 20) || q >= (1u << 11))
                throw std::runtime_error("hop table index exceeds the packed entry range");
            h.ent.push_back((uint32_t)i | ((uint32_t)q << 20) | ((uint32_t)(e & 1) << 31));
        }
        h.rowptr[f + 1] = (int64_t)h.ent.size();
    }
    return h;
}

// operators of a list grouped by their annihilated pair (c | d): an SD can only be
// acted on by the operators of the pairs it occupies
using OpsByPair = std::unordered_map<uint64_t, std::vector<int32_t>>;
static OpsByPair group_by_pair(const OpList& L) {
    OpsByPair g;
    for (size_t q = 0; q < L.ops.size(); q++)
        if (L.V[q] != 0.0) g[L.ops[q].c | L.ops[q].d].push_back((int32_t)q);
    return g;
}

static PairSym build_pair_matrix(const Species& s, int set, const OpList& L, const OpsByPair& byp) {
    const auto& S = s.sets[set];
    size_t n = S.size();
    std::vector<std::vector<std::pair<int32_t, double>>> rows(n);
    for (size_t i = 0; i < n; i++) {
        const uint64_t sd = S[i];
        for (uint64_t r1 = sd; r1; r1 &= r1 - 1) {
            const uint64_t b1 = r1 & (~r1 + 1);
            for (uint64_t r2 = r1 & (r1 - 1); r2; r2 &= r2 - 1) {
                const uint64_t b2 = r2 & (~r2 + 1);
                auto it = byp.find(b1 | b2);
                if (it == byp.end()) continue;
                for (int32_t q : it->second) {
                    uint64_t fsd; int neg;
                    if (!apply2(sd, L.ops[q], &fsd, &neg)) continue;
                    int64_t f = s.find(set, fsd);
                    if (f < 0 || (int64_t)i > f) continue;          // entries kept only for i <= f
                    double coef = L.V[q] * (neg ? -1.0 : 1.0) * ((int64_t)i == f ? 0.5 : 1.0);
                    // the reference applies every entry in both directions
                    rows[f].push_back({(int32_t)i, coef});
                    rows[i].push_back({(int32_t)f, coef});
                }
            }
        }
    }
    PairSym m;
    m.rowptr.assign(n + 1, 0);
    for (size_t r = 0; r < n; r++) {
        auto& R = rows[r];
        std::stable_sort(R.begin(), R.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (size_t k = 0; k < R.size();) {
            int32_t c = R[k].first;
            double v = 0.0;
            for (; k < R.size() && R[k].first == c; k++) v += R[k].second;
            m.col.push_back(c);
            m.val.push_back(v);
        }
        m.rowptr[r + 1] = (int64_t)m.col.size();
    }
    return m;
}

void build_coefficients(const Snt& snt, const Basis& b, Tables& t, int nthreads) {
    const Species& P = b.sp[0];
    const Species& N = b.sp[1];
    // one-body energies (sum over orbits in orbit order, as dot(SPEs, jocc))
    t.ep.assign(P.sets.size(), {});
    t.en.assign(N.sets.size(), {});
    for (size_t k = 0; k < P.sets.size(); k++) {
        t.ep[k].resize(P.sets[k].size());
        for (size_t i = 0; i < P.sets[k].size(); i++) {
            double e = 0.0;
            for (int o = 0; o < P.norb; o++) e += snt.spe_p[o] * P.jocc[k][i * P.norb + o];
            t.ep[k][i] = e;
        }
    }
    for (size_t k = 0; k < N.sets.size(); k++) {
        t.en[k].resize(N.sets[k].size());
        for (size_t i = 0; i < N.sets[k].size(); i++) {
            double e = 0.0;
            for (int o = 0; o < N.norb; o++) e += snt.spe_n[o] * N.jocc[k][i * N.norb + o];
            t.en[k][i] = e;
        }
    }
    OpList pp = t1_ops(snt, P, 0), nn = t1_ops(snt, N, 1);
    t.n_pp_ops = (int64_t)pp.ops.size();
    t.n_nn_ops = (int64_t)nn.ops.size();
    t.pp.assign(P.sets.size(), {});
    t.nn.assign(N.sets.size(), {});
    // only M-sets that appear in some block are needed
    std::vector<char> pneed(P.sets.size(), 0), nneed(N.sets.size(), 0);
    for (int k = 0; k < b.nb; k++) { pneed[b.pset[k]] = 1; nneed[b.nset[k]] = 1; }
    int64_t np = (int64_t)P.sets.size();
    const OpsByPair ppg = group_by_pair(pp), nng = group_by_pair(nn);
    parallel_for(np + (int64_t)N.sets.size(), nthreads, [&](int64_t q) {
        if (q < np) { if (pneed[q]) t.pp[q] = build_pair_matrix(P, (int)q, pp, ppg); }
        else { int64_t r = q - np; if (nneed[r]) t.nn[r] = build_pair_matrix(N, (int)r, nn, nng); }
    });
    // V_eff
    auto pn = pn_ops(snt, P, N);
    int64_t npn = 0;
    for (auto& kv : pn) npn += (int64_t)kv.second.ops.size();
    t.n_pn_ops = npn;
    const OneBodyOps& obp = t.ob[0];
    const OneBodyOps& obn = t.ob[1];
    size_t nd = obp.dMs.size();
    t.veff.assign(nd, {});
    t.veff_rows.assign(nd, 0);
    t.veff_cols.assign(nd, 0);
    for (size_t k = 0; k < nd; k++) {
        int dM = obp.dMs[k];
        int kn = obn.index_of_dM(-dM);
        int rows = (int)obp.ops[k].size();
        int cols = kn >= 0 ? (int)obn.ops[kn].size() : 0;
        t.veff_rows[k] = rows;
        t.veff_cols[k] = cols;
        t.veff[k].assign((size_t)rows * cols, 0.0);
        if (cols == 0) continue;
        // lookup of the reference operator coefficient:
        //   dM > 0 : forward operator (a', b', c', d')             (pn2 of operate_H!)
        //   dM <= 0: transposed operator (c', d', a', b') of -dM   (pn1 of operate_H!)
        int src = dM > 0 ? dM : -dM;
        auto it = pn.find(src);
        if (it == pn.end()) continue;
        const OpList& L = it->second;
        for (int r = 0; r < rows; r++) {
            uint64_t pa = bit_of(P, obp.ops[k][r].first), pc = bit_of(P, obp.ops[k][r].second);
            for (int c = 0; c < cols; c++) {
                uint64_t nb_ = bit_of(N, obn.ops[kn][c].first), nd_ = bit_of(N, obn.ops[kn][c].second);
                std::array<uint64_t, 4> key = dM > 0 ? std::array<uint64_t, 4>{pa, nb_, pc, nd_}
                                                     : std::array<uint64_t, 4>{pc, nd_, pa, nb_};
                auto f = L.seen.find(key);
                if (f != L.seen.end()) t.veff[k][(size_t)r * cols + c] = L.V[f->second];
            }
        }
    }
}

Tables build_tables(const Snt& snt, const Basis& b, int nthreads) {
    Tables t;
    build_onebody_ops(b.sp[0], t.ob[0]);
    build_onebody_ops(b.sp[1], t.ob[1]);
    // ordered pairs of blocks (output bf, input bi) coupled by a proton one-body dM
    std::vector<std::array<int, 3>> want;   // (species, iset, fset)
    for (int bf = 0; bf < b.nb; bf++)
        for (int bi = 0; bi < b.nb; bi++) {
            int dM = b.Mp[bf] - b.Mp[bi];
            if (t.ob[0].index_of_dM(dM) < 0 || t.ob[1].index_of_dM(-dM) < 0) continue;
            want.push_back({0, b.pset[bi], b.pset[bf]});
            want.push_back({1, b.nset[bi], b.nset[bf]});
        }
    std::sort(want.begin(), want.end());
    want.erase(std::unique(want.begin(), want.end()), want.end());
    t.hops.resize(want.size());
    parallel_for((int64_t)want.size(), nthreads, [&](int64_t q) {
        const auto& w = want[q];
        t.hops[q] = build_hops(b.sp[w[0]], t.ob[w[0]], w[0], w[1], w[2]);
    });
    for (size_t q = 0; q < want.size(); q++) t.hop_index[want[q]] = (int)q;
    build_coefficients(snt, b, t, nthreads);
    return t;
}

// ------------------------------------------------------------------ moments
double wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
    if (m1 + m2 + m3 != 0 || ((j1 - j2 - m3) & 1)) return 0.0;
    const int e = (j1 - j2 - m3) / 2;
    return ((e & 1) ? -1.0 : 1.0) / std::sqrt((double)(j3 + 1)) * clebsch_gordan(j1, m1, j2, m2, j3, -m3);
}

// Delta(abc)^2 of doubled arguments; false if the triangle or the parity fails
static bool delta2(int a, int b, int c, __float128* d) {
    if (((a + b + c) & 1) || a + b < c || a + c < b || b + c < a) return false;
    *d = qfact((a + b - c) / 2) * qfact((a - b + c) / 2) * qfact((-a + b + c) / 2) / qfact((a + b + c) / 2 + 1);
    return true;
}

// Racah's formula in quad precision, rounded once to double
double wigner6j(int j1, int j2, int j3, int j4, int j5, int j6) {
    __float128 d[4];
    if (!delta2(j1, j2, j3, &d[0]) || !delta2(j1, j5, j6, &d[1]) || !delta2(j4, j2, j6, &d[2]) ||
        !delta2(j4, j5, j3, &d[3]))
        return 0.0;
    const int a[4] = {(j1 + j2 + j3) / 2, (j1 + j5 + j6) / 2, (j4 + j2 + j6) / 2, (j4 + j5 + j3) / 2};
    const int c[3] = {(j1 + j2 + j4 + j5) / 2, (j2 + j3 + j5 + j6) / 2, (j3 + j1 + j6 + j4) / 2};
    const int t0 = std::max(std::max(a[0], a[1]), std::max(a[2], a[3]));
    const int t1 = std::min(std::min(c[0], c[1]), c[2]);
    __float128 s = 0;
    for (int t = t0; t <= t1; t++) {
        __float128 den = 1;
        for (int x : a) den *= qfact(t - x);
        for (int x : c) den *= qfact(x - t);
        s += ((t & 1) ? -1 : 1) * qfact(t + 1) / den;
    }
    if (s == 0) return 0.0;
    return (double)(sqrtq(d[0] * d[1] * d[2] * d[3]) * s);
}

static __float128 dfact(int n) {
    __float128 r = 1;
    for (; n > 1; n -= 2) r *= n;
    return r;
}

// binomial coefficient with Julia's extension to negative n
static __float128 binom(int n, int k) {
    if (k < 0) return 0;
    if (n < 0) return ((k & 1) ? -1 : 1) * binom(k - n - 1, k);
    if (k > n) return 0;
    __float128 r = 1;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

double radius_power(int k, int n1, int l1, int n2, int l2, double bpar) {
    const int ll = l1 + l2 + k;
    const int ll1 = (l2 - l1 + k) / 2, ll2 = (l1 - l2 + k) / 2;   // truncation toward zero
    __float128 s = 0;
    for (int i = std::max({0, n1 - ll1, n2 - ll2}); i <= std::min(n1, n2); i++)
        s += binom(ll1, n1 - i) * binom(ll2, n2 - i) * dfact(ll + 2 * i + 1) / dfact(2 * i);
    s *= sqrtq(dfact(2 * n1) / dfact(2 * n1 + 2 * l1 + 1));
    s *= sqrtq(dfact(2 * n2) / dfact(2 * n2 + 2 * l2 + 1));
    return (double)s / std::sqrt(std::pow(2.0, k)) * (((n1 - n2) & 1) ? -1.0 : 1.0) * std::pow(bpar, k);
}

double ho_bpar_main_sm(int A) {
    const double amass = (938.27231 + 939.56563) / 2, hc = 197.32696;
    const double a = (double)A;
    const double hw = (16 <= A && A <= 40) ? 45.0 * std::pow(a, -1.0 / 3.0) - 25.0 * std::pow(a, -2.0 / 3.0)
                                           : 41.0 * std::pow(a, -1.0 / 3.0);
    return hc / std::sqrt(hw * amass);
}

MomentOps build_moment_ops(const Snt& snt, const Basis& b, double bpar, const double* g, const double* ec) {
    MomentOps mo;
    mo.bpar = bpar;
    const double c_mu = std::sqrt(3.0 / (4.0 * M_PI)), c_Q = std::sqrt(5.0 / (16.0 * M_PI));
    for (int sp = 0; sp < 2; sp++) {
        const Species& S = b.sp[sp];
        const std::vector<Orbit>& sps = sp == 0 ? snt.p_sps : snt.n_sps;
        const double gl = sp == 0 ? g[0] : g[1], gs = sp == 0 ? g[2] : g[3], e = ec[sp];
        // the M-sets of this species that the basis uses
        std::vector<char> used(S.sets.size(), 0);
        for (int k = 0; k < b.nb; k++) used[sp == 0 ? b.pset[k] : b.nset[k]] = 1;
        // one-body terms (m-state a <- m-state c, coefficient) of each operator
        std::vector<std::array<int, 2>> terms[2];
        std::vector<double> coef[2];
        for (size_t i = 0; i < sps.size(); i++)
            for (size_t f = i; f < sps.size(); f++) {
                const Orbit oi = sps[i], of = sps[f];
                const int ji = oi.j2, jf = of.j2, li = oi.l, lf = of.l;
                const double ex_r2 = radius_power(2, oi.n, li, of.n, lf, bpar);
                const double l6j = wigner6j(2 * lf, 2 * li, 2, ji, jf, 1);
                const double s6j = wigner6j(1, 1, 2, ji, jf, 2 * lf);
                const double Q3j = wigner3j(ji, 4, jf, 1, 0, -1);
                const double pfac = std::sqrt((double)((ji + 1) * (jf + 1)));
                const std::vector<int> kc = mstates_of(S, (int)i), ka = mstates_of(S, (int)f);
                for (int mode = 0; mode < 2; mode++) {
                    if (mode == 0 && (oi.n != of.n || li != lf)) continue;
                    for (int c : kc)
                        for (int a : ka) {
                            const int m = S.m2[c];
                            if (S.m2[a] != m) continue;
                            double v;
                            if (mode == 0) {
                                const int s1 = (jf - m) / 2 + lf + (ji + 3) / 2, s2 = (jf - m) / 2 + lf + (jf + 3) / 2;
                                const double fac1 = ((s1 & 1) ? -1.0 : 1.0) * l6j * std::sqrt((double)(lf * (lf + 1) * (2 * lf + 1)));
                                const double fac2 = ((s2 & 1) ? -1.0 : 1.0) * s6j * std::sqrt(1.5);
                                v = pfac * wigner3j(jf, 2, ji, -m, 0, m) * (fac1 * gl + fac2 * gs) * c_mu;
                            } else {
                                if ((li + lf) & 1) continue;      // (1 + (-1)^(li+lf)) = 0
                                const int sq = jf + (1 - m) / 2;
                                v = ((sq & 1) ? -2.0 : 2.0) * wigner3j(jf, 4, ji, -m, 0, m) * Q3j * pfac * e * c_Q * ex_r2;
                            }
                            if (v == 0.0) continue;
                            terms[mode].push_back({a, c});
                            coef[mode].push_back(v);
                        }
                }
            }
        for (int mode = 0; mode < 2; mode++) {
            std::vector<PairSym>& out = mo.op[mode][sp];
            out.assign(S.sets.size(), PairSym{});
            for (size_t set = 0; set < S.sets.size(); set++) {
                const auto& F = S.sets[set];
                PairSym& P = out[set];
                P.rowptr.assign(F.size() + 1, 0);
                if (!used[set]) continue;
                std::vector<std::map<int32_t, double>> rows(F.size());
                for (size_t q = 0; q < terms[mode].size(); q++) {
                    const uint64_t ba = bit_of(S, terms[mode][q][0]), bc = bit_of(S, terms[mode][q][1]);
                    for (size_t ni = 0; ni < F.size(); ni++) {
                        const uint64_t phi = F[ni];
                        if (!(phi & bc)) continue;
                        const uint64_t mid = phi & ~bc;
                        if (mid & ba) continue;
                        const int64_t nf = S.find((int)set, mid | ba);
                        if (nf < 0) continue;
                        // fermion sign of a+ c on phi (calc_phase_1! with its leading 1)
                        const int ex = 1 + popc(phi & at_or_above(bc)) + popc(mid & at_or_above(ba));
                        double w = ((ex & 1) ? -1.0 : 1.0) * coef[mode][q] * ((int64_t)ni == nf ? 0.5 : 1.0);
                        rows[nf][(int32_t)ni] += w;
                        rows[ni][(int32_t)nf] += w;
                    }
                }
                for (size_t r = 0; r < F.size(); r++) {
                    for (auto& kv : rows[r]) { P.col.push_back(kv.first); P.val.push_back(kv.second); }
                    P.rowptr[r + 1] = (int64_t)P.col.size();
                }
            }
        }
    }
    return mo;
}

TransitionOps build_transition_ops(const Snt& snt, const Basis& bl, const Basis& br, double bpar, const double* g,
                                   const double* ec) {
    TransitionOps to;
    to.bpar = bpar;
    to.mu2 = bl.twoM - br.twoM;
    if (to.mu2 < 0 || (to.mu2 & 1))
        throw std::invalid_argument("the left basis needs 2M >= the right basis' 2M, of the same parity");
    const int omu = to.mu2 / 2;
    if (omu > 2) throw std::invalid_argument("only M1 and E2 transitions: |M_left - M_right| <= 2");
    for (int sp = 0; sp < 2; sp++)
        if (bl.sp[sp].Ms != br.sp[sp].Ms || bl.sp[sp].nv != br.sp[sp].nv || bl.sp[sp].L != br.sp[sp].L)
            throw std::invalid_argument("left and right bases of different nuclei or spaces");
    const double c_mu = std::sqrt(3.0 / (4.0 * M_PI)), c_Q = std::sqrt(5.0 / (16.0 * M_PI));
    // right blocks by (proton set, neutron set)
    std::map<std::pair<int, int>, int> rblock;
    for (int k = 0; k < br.nb; k++) rblock[{br.pset[k], br.nset[k]}] = k;
    for (int sp = 0; sp < 2; sp++) {
        const Species& S = bl.sp[sp];
        const std::vector<Orbit>& sps = sp == 0 ? snt.p_sps : snt.n_sps;
        const double gl = sp == 0 ? g[0] : g[1], gs = sp == 0 ? g[2] : g[3], e = ec[sp];
        // left M-sets of this species whose source block exists in the right basis
        std::vector<char> used(S.sets.size(), 0);
        for (int k = 0; k < bl.nb; k++) {
            const int t = sp == 0 ? bl.pset[k] : bl.nset[k];
            const int src = S.set_of_M(S.Ms[t] - to.mu2);
            if (src < 0) continue;
            const std::pair<int, int> key = sp == 0 ? std::make_pair(src, bl.nset[k]) : std::make_pair(bl.pset[k], src);
            if (rblock.count(key)) used[t] = 1;
        }
        // one-body terms (m-state a <- m-state c, coefficient), m_a - m_c = mu2
        std::vector<std::array<int, 2>> terms[2];
        std::vector<double> coef[2];
        for (size_t i = 0; i < sps.size(); i++)
            for (size_t f = 0; f < sps.size(); f++) {
                const Orbit oi = sps[i], of = sps[f];
                const int ji = oi.j2, jf = of.j2, li = oi.l, lf = of.l;
                const double ex_r2 = radius_power(2, oi.n, li, of.n, lf, bpar);
                const double l6j = wigner6j(2 * lf, 2 * li, 2, ji, jf, 1);
                const double s6j = wigner6j(1, 1, 2, ji, jf, 2 * lf);
                const double Q3j = wigner3j(ji, 4, jf, 1, 0, -1);
                const double pfac = std::sqrt((double)((ji + 1) * (jf + 1)));
                const std::vector<int> kc = mstates_of(S, (int)i), ka = mstates_of(S, (int)f);
                for (int mode = 0; mode < 2; mode++) {
                    if (mode == 0 && (oi.n != of.n || li != lf || omu > 1)) continue;
                    for (int c : kc)
                        for (int a : ka) {
                            const int mc = S.m2[c], ma = S.m2[a];
                            if (ma - mc != to.mu2) continue;
                            double v;
                            if (mode == 0) {
                                const int s1 = (jf - ma) / 2 + lf + (ji + 3) / 2, s2 = (jf - ma) / 2 + lf + (jf + 3) / 2;
                                const double fac1 = ((s1 & 1) ? -1.0 : 1.0) * l6j * std::sqrt((double)(lf * (lf + 1) * (2 * lf + 1)));
                                const double fac2 = ((s2 & 1) ? -1.0 : 1.0) * s6j * std::sqrt(1.5);
                                v = pfac * wigner3j(jf, 2, ji, -ma, to.mu2, mc) * (fac1 * gl + fac2 * gs) * c_mu;
                            } else {
                                if ((li + lf) & 1) continue;      // (1 + (-1)^(li+lf)) = 0
                                const int sq = jf + (1 - ma) / 2;
                                v = ((sq & 1) ? -2.0 : 2.0) * wigner3j(jf, 4, ji, -ma, to.mu2, mc) * Q3j * pfac * e * c_Q *
                                    ex_r2;
                            }
                            if (v == 0.0) continue;
                            terms[mode].push_back({a, c});
                            coef[mode].push_back(v);
                        }
                }
            }
        for (int mode = 0; mode < 2; mode++) {
            std::vector<PairSym>& out = to.op[mode][sp];
            out.assign(S.sets.size(), PairSym{});
            for (size_t t = 0; t < S.sets.size(); t++) {
                const auto& F = S.sets[t];
                PairSym& P = out[t];
                P.rowptr.assign(F.size() + 1, 0);
                if (!used[t]) continue;
                const int src = S.set_of_M(S.Ms[t] - to.mu2);
                const auto& I = S.sets[src];
                std::vector<std::map<int32_t, double>> rows(F.size());
                for (size_t q = 0; q < terms[mode].size(); q++) {
                    const uint64_t ba = bit_of(S, terms[mode][q][0]), bc = bit_of(S, terms[mode][q][1]);
                    for (size_t ni = 0; ni < I.size(); ni++) {
                        const uint64_t phi = I[ni];
                        if (!(phi & bc)) continue;
                        const uint64_t mid = phi & ~bc;
                        if (mid & ba) continue;
                        const int64_t nf = S.find((int)t, mid | ba);
                        if (nf < 0) continue;
                        // fermion sign of a+ c on phi (calc_phase_1! with its leading 1)
                        const int ex = 1 + popc(phi & at_or_above(bc)) + popc(mid & at_or_above(ba));
                        rows[nf][(int32_t)ni] += ((ex & 1) ? -1.0 : 1.0) * coef[mode][q];
                    }
                }
                for (size_t r = 0; r < F.size(); r++) {
                    for (auto& kv : rows[r]) { P.col.push_back(kv.first); P.val.push_back(kv.second); }
                    P.rowptr[r + 1] = (int64_t)P.col.size();
                }
            }
        }
    }
    return to;
}

}  // namespace smgpu
