// Host-side EC-FDTD operator builder for box scenes (see include/ems_operator.h).
//
// Every floating-point expression keeps the reference's operand order and
// precision (float EC arrays, double intermediates), and the file is compiled
// with -ffp-contract=off, so the float32 coefficients come out bit-identical
// to the reference Operator for the same mesh and materials.
//
// Reference functions restated here (FDTD/operator.cpp unless noted):
//   GetDiscLine/GetDiscDelta/GetEdgeLength      :144-213
//   SnapToMeshLine/SnapToMesh/SnapBox2Mesh        :258-327
//   Calc_ECOperatorPos                            :956-984
//   CalcECOperator                                :986-1097
//   ApplyElectricBC / ApplyMagneticBC             :1099-1187
//   Calc_ECPos                                    :1189-1254
//   GetRawDiscDelta                               :1256-1266
//   AverageMatCellCenter / AverageMatQuarterCell  :1336-1564
//   Calc_LumpedElements (parallel RC)             :1586-1763
//   CalcTimestep_Var1 / Var3                      :1892-2030
//   CalcPEC_Range                                 :2045-2083
//   Excitation signals                            FDTD/excitation.cpp:150-276
//   Operator_Ext_Excitation::BuildExtension       FDTD/extensions/operator_ext_excitation.cpp:105-297
//   Operator_Ext_UPML::Create_UPML/CalcGradingKappa/BuildExtension
//                                                 FDTD/extensions/operator_ext_upml.cpp:69-445
//   Operator_Ext_Mur_ABC::BuildExtension          FDTD/extensions/operator_ext_mur_abc.cpp:80-186
//   Engine_Ext_Mur_ABC late start                 FDTD/extensions/engine_ext_mur_abc.cpp:28-60
//   Operator_Ext_LorentzMaterial::BuildExtension  FDTD/extensions/operator_ext_lorentzmaterial.cpp:120-445
//   Operator_Ext_LumpedRLC::BuildExtension        FDTD/extensions/operator_ext_lumpedRLC.cpp:112-548
//
// Material lookup: the CSXCAD rule "highest priority primitive containing the
// point wins" is evaluated with per-axis bitmasks over boxes ranked by
// (priority, earlier insertion first): the winner at a point is the highest
// rank set in mask_x & mask_y & mask_z.
#include "ems_operator.h"

#include <algorithm>
#include <array>
#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <sys/mman.h>
#include <new>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr double kEPS0 = 8.85418781762e-12;  // tools/constants.h:23
constexpr double kMUE0 = 1.256637062e-6;     // tools/constants.h:24
constexpr double kC0 = 299792458;            // tools/constants.h:25
constexpr double kZ0 = 376.730313461;        // tools/constants.h:26
constexpr double kPI = 3.141592653589793238462643383279;  // tools/constants.h:27

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

// Point_Line_Distance of CSXCAD (Cartesian mesh): foot = (P - s).(e - s) / |e - s|^2
// is the position of the foot of the perpendicular along s -> e, dist the
// distance of P from the line.  A zero-length line gives foot = NaN.
static void point_line_distance(const double* P, const double* s, const double* e, double& foot, double& dist)
{
    const double d[3] = {e[0] - s[0], e[1] - s[1], e[2] - s[2]};
    const double LL = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    foot = ((P[0] - s[0]) * d[0] + (P[1] - s[1]) * d[1] + (P[2] - s[2]) * d[2]) / LL;
    const double q[3] = {P[0] - (s[0] + foot * d[0]), P[1] - (s[1] + foot * d[1]), P[2] - (s[2] + foot * d[2])};
    dist = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
}

// A primitive: an axis-aligned box (lo, hi), or a cylinder (CSPrimCylinder:
// axis ax0 -> ax1, radius rad) or cylindrical shell (CSPrimCylindricalShell:
// mid radius rad, wall width) whose bounding box is lo, hi.  A point is inside
// a cylinder when the foot of its perpendicular on the axis lies within the
// axis segment (0 <= foot <= 1) and its distance from the axis is at most the
// radius; inside a shell when that distance is within width / 2 of rad.
struct Box {
    double lo[3], hi[3];
    int prio;
    int order;
    int shape = 0;            // 0 box, 1 cylinder, 2 cylindrical shell
    double ax0[3] = {0, 0, 0}, ax1[3] = {0, 0, 0};
    double rad = 0, width = 0;
    bool contains(const double* P) const
    {
        if (shape == 0) return true;   // the per-axis masks decide for boxes
        double foot, dist;
        point_line_distance(P, ax0, ax1, foot, dist);
        if (!(foot >= 0 && foot <= 1)) return false;
        if (shape == 1) return dist <= rad;
        return std::fabs(dist - rad) <= width / 2;
    }
    void set_cylinder(const double* a, const double* b, double r, double w)
    {
        shape = w > 0 ? 2 : 1;
        rad = r;
        width = w > 0 ? w : 0;
        const double reach = r + width / 2;
        for (int n = 0; n < 3; ++n) {
            ax0[n] = a[n];
            ax1[n] = b[n];
            lo[n] = std::min(a[n], b[n]) - reach;
            hi[n] = std::max(a[n], b[n]) + reach;
        }
    }
};

struct MatBox : Box {
    double eps[3], mue[3], kappa[3], sigma[3];
    bool metal;
    int grid = -1;   // >= 0: per-cell eps_r / kappa from emsop::grids[grid]
    int disp = -1;   // >= 0: Lorentz / Drude dispersion emsop::disps[disp]
};

// Lorentz / Drude dispersion of a material (CSPropLorentzMaterial): per
// order o and direction n, p[(o * 6 + k) * 3 + n] with k = eps plasma
// frequency (Hz), eps relaxation time (s), eps Lorentz pole frequency (Hz),
// then the same three for mue.
struct Dispersion {
    unsigned norder;
    bool debye = false;   // CSPropDebyeMaterial: k = 0 eps delta, 1 eps relaxation time (s)
    std::vector<double> p;
    double at(unsigned o, int k, int n) const { return o < norder ? p[((size_t)o * 6 + k) * 3 + n] : 0.0; }
};

// Operator_Ext_LorentzMaterial tables of one dispersion order
struct LorOrder {
    bool von = false, ion = false, vlon = false, ilon = false;
    std::vector<unsigned> pos[3];
    std::vector<float> t[6][3];     // v_int, v_ext, i_int, i_ext, v_lor, i_lor
};

// Operator_Ext_LumpedRLC tables
struct RlcTab {
    std::vector<unsigned> pos[3];
    std::vector<int> dir;
    std::vector<float> c[9];        // i2v, ilv, vvd, vv2, vj1, vj2, ib0, b1, b2
};

// Material grid: a box of whole mesh cells with its own eps_r and kappa per
// cell (isotropic), equivalent to one-cell boxes of one priority added in
// x-fastest order: a point on a face shared by grid cells takes the cell of
// the lowest index along each axis.
struct MatGrid {
    unsigned i0[3], n[3];
    std::vector<double> eps, kappa;   // [n2][n1][n0] (x fastest)
};

struct LumpedBox : Box {
    int dir;
    double R, C, L;
    bool caps, parallel;
};

// Excitation weight functions (CSPropExcitation::SetWeightFunction): one
// expression per direction in x, y, z, rho, r, a (drawing units, a = the
// azimuth); the amplitude of an edge is weight(Yee point) * value * length
struct Expr;
struct ExcWeight {
    std::shared_ptr<const Expr> f[3];
};
double weight_eval(const ExcWeight& w, int n, const double* P);
inline double weighted(const std::shared_ptr<const ExcWeight>& w, int n, double val, const double* P)
{
    return w ? weight_eval(*w, n, P) * val : val;
}

struct ExcBox : Box {
    int type;
    double val[3];
    double delay;
    std::shared_ptr<const ExcWeight> weight;
};

struct Slab {
    unsigned start[3], num[3];
    std::vector<float> a[6];  // vv, vvfo, vvfn, ii, iifo, iifn
};

struct MurFace {
    int ny;
    unsigned line, shift, start_ts;
    std::vector<float> cP, cPP;
};

// Polyline primitive (CSPrimCurve) treated as a wire: a metal curve zeroes
// vv / vi on the mesh edges of its path, an excitation curve puts its edges on
// the voltage excitation list (points in drawing units, in the order added)
struct Curve {
    std::vector<double> pts;   // [npts][3]
    bool exc = false;
    int type = 0;              // excitation type (0 soft, 1 hard)
    double val[3] = {0, 0, 0};
    double delay = 0;
    std::shared_ptr<const ExcWeight> weight;
};

// one mesh edge of a path: component dir at node pos
struct PathEdge {
    unsigned pos[3];
    int dir;
};

struct ExcList {
    std::vector<unsigned> idx[3];
    std::vector<unsigned short> dir;
    std::vector<float> amp;
    std::vector<unsigned> delay;
};

void parallel_for(unsigned n, int nthreads, const std::function<void(unsigned, unsigned)>& fn)
{
    unsigned nt = nthreads > 0 ? (unsigned)nthreads : std::max(1u, std::thread::hardware_concurrency());
    nt = std::min(nt, std::max(1u, n));
    if (nt <= 1) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> th;
    unsigned per = n / nt, rem = n % nt, s = 0;
    for (unsigned t = 0; t < nt; ++t) {
        unsigned c = per + (t < rem ? 1 : 0);
        th.emplace_back(fn, s, s + c);
        s += c;
    }
    for (auto& t : th) t.join();
}

// uninitialised array: the big operator arrays are first touched by the
// threads that fill them instead of a serial zero fill
template <typename T>
struct PodVec {
    struct Free {
        void operator()(T* q) const { std::free(q); }
    };
    std::unique_ptr<T[], Free> p;
    size_t n = 0;
    void alloc(size_t count)
    {
        // 2 MiB aligned and backed by transparent huge pages where the kernel allows it
        const size_t align = size_t(2) << 20;
        const size_t bytes = (std::max<size_t>(count, 1) * sizeof(T) + align - 1) / align * align;
        void* q = nullptr;
        if (posix_memalign(&q, align, bytes) != 0) throw std::bad_alloc();
        madvise(q, bytes, MADV_HUGEPAGE);
        p.reset(static_cast<T*>(q));
        n = count;
    }
    void clear()
    {
        p.reset();
        n = 0;
    }
    T& operator[](size_t i) { return p[i]; }
    const T& operator[](size_t i) const { return p[i]; }
    T* data() { return p.get(); }
    const T* data() const { return p.get(); }
    size_t size() const { return n; }
    bool empty() const { return n == 0; }
};

// ranked box set with per-axis membership masks
template <typename B>
struct Ranked {
    std::vector<const B*> box;  // ascending rank
    unsigned W = 0;             // 64-bit words per mask
    void init(const std::vector<B>& all, std::function<bool(const B&)> keep)
    {
        box.clear();
        for (const auto& b : all)
            if (keep(b)) box.push_back(&b);
        std::stable_sort(box.begin(), box.end(), [](const B* a, const B* b) {
            if (a->prio != b->prio) return a->prio < b->prio;
            return a->order > b->order;  // earlier insertion ranks higher among equals
        });
        W = (unsigned)((box.size() + 63) / 64);
    }
    // mask[i*W + w] for coordinates c[i] along axis n
    std::vector<uint64_t> masks(int n, const std::vector<double>& c) const
    {
        std::vector<uint64_t> m((size_t)c.size() * W, 0);
        for (size_t r = 0; r < box.size(); ++r) {
            double a = std::min(box[r]->lo[n], box[r]->hi[n]);
            double b = std::max(box[r]->lo[n], box[r]->hi[n]);
            for (size_t i = 0; i < c.size(); ++i)
                if (c[i] >= a && c[i] <= b) m[i * W + r / 64] |= (uint64_t)1 << (r % 64);
        }
        return m;
    }
    // highest rank set in the AND of three masks, -1 if none
    inline int winner(const uint64_t* a, const uint64_t* b, const uint64_t* c) const
    {
        for (int w = (int)W - 1; w >= 0; --w) {
            uint64_t x = a[w] & b[w] & c[w];
            if (x) return w * 64 + 63 - __builtin_clzll(x);
        }
        return -1;
    }
    // as winner, for sets that may hold cylinders: a set bit of a cylinder
    // only means P is in its bounding box, so it wins only if it contains P
    inline int winner_at(const uint64_t* a, const uint64_t* b, const uint64_t* c, const double* P) const
    {
        for (int w = (int)W - 1; w >= 0; --w) {
            uint64_t x = a[w] & b[w] & c[w];
            while (x) {
                const int bit = 63 - __builtin_clzll(x);
                const int r = w * 64 + bit;
                if (box[r]->contains(P)) return r;
                x &= ~((uint64_t)1 << bit);
            }
        }
        return -1;
    }
};

}  // namespace

struct emsop {
    // scene
    std::vector<double> lines[3];
    unsigned N[3] = {0, 0, 0};
    double gd = 1;
    double bg[4] = {1, 1, 0, 0};
    std::vector<MatBox> mats;
    std::vector<MatGrid> grids;
    std::vector<LumpedBox> lumped;
    std::vector<ExcBox> excs;
    std::vector<Curve> curves;
    int last_exc = 0;        // 1: the excitation primitive added last is excs.back(), 2: curves.back()
    int bc[6] = {0, 0, 0, 0, 0, 0};
    unsigned pml[6] = {8, 8, 8, 8, 8, 8};
    double vph[6] = {0, 0, 0, 0, 0, 0};
    int sigType = -1;
    double f0 = 0, fc = 0;
    std::string sigFunc;     // custom signal: f(t) in the function-parser syntax
    int tsMethod = 3;
    double forcedDt = 0, tsFactor = 1;
    bool cellConst = false;
    int nthreads = 0;
    bool built = false;
    // derived mesh tables
    std::vector<double> primD[3], dualL[3], dualD[3];
    // results
    double dT = 0;
    unsigned nyquist = 0;
    double period = 0;
    unsigned sigLen = 0;
    std::vector<float> sigV, sigC;
    PodVec<float> coeff[4];  // vv vi ii iv
    PodVec<float> ec[4];     // C G L R
    // material sampling tables, one set per field component
    Ranked<MatBox> matrk;
    struct Samp {
        std::vector<uint64_t> eps_n, P_up, P_dn, PP_up, PP_dn, mu_dn, mu_up, P_half, PP_half, dual_n, dual_P, dual_PP;
        // sample coordinates (for material grids)
        std::vector<double> c_eps_n, cPu, cPd, cPh, cPPu, cPPd, cPPh, c_mu_dn, c_mu_up;
    } samp[3];
    ExcList exc[2];
    std::vector<Slab> upml;
    std::vector<MurFace> mur;
    std::vector<Dispersion> disps;
    std::vector<LorOrder> lor;
    RlcTab rlc;

    size_t size() const { return (size_t)N[0] * N[1] * N[2]; }
    size_t idx(int n, unsigned x, unsigned y, unsigned z) const
    {
        return (((size_t)n * N[0] + x) * N[1] + y) * N[2] + z;
    }
    // GetDiscDelta(n,i,false)*gridDelta / GetDiscDelta(n,i,true)*gridDelta
    double EL(int n, unsigned i, bool dual) const { return (dual ? dualD[n][i] : primD[n][i]) * gd; }
    // GetNodeWidth(n, loc, true) for an int index that may be negative
    double Wp(int n, int i) const { return i < 0 ? 0.0 : primD[n][i] * gd; }
    // GetRawDiscDelta (operator.cpp:1256-1266)
    double rawD(int n, int pos) const
    {
        const auto& l = lines[n];
        if (pos < 0) return l[0] - l[1];
        if (pos >= (int)N[n] - 1) return l[N[n] - 2] - l[N[n] - 1];
        return l[pos + 1] - l[pos];
    }

    void mesh_tables()
    {
        for (int n = 0; n < 3; ++n) {
            const auto& l = lines[n];
            unsigned c = N[n];
            primD[n].assign(c, 0);
            dualL[n].assign(c, 0);
            dualD[n].assign(c, 0);
            for (unsigned i = 0; i < c; ++i) {
                primD[n][i] = (i < c - 1) ? l[i + 1] - l[i] : l[i] - l[i - 1];
                dualL[n][i] = (i < c - 1) ? 0.5 * (l[i] + l[i + 1]) : l[i] + 0.5 * (l[i] - l[i - 1]);
            }
            for (unsigned i = 0; i < c; ++i) dualD[n][i] = (i > 0) ? dualL[n][i] - dualL[n][i - 1] : l[1] - l[0];
        }
    }

    unsigned snap_line(int n, double coord, bool& inside, bool dual) const
    {
        // SnapToMeshLine (operator.cpp:258-287)
        inside = false;
        const auto& l = lines[n];
        if (coord < l[0]) return 0;
        if (coord > l[N[n] - 1]) return N[n] - 1;
        inside = true;
        if (!dual) {
            for (unsigned i = 0; i < N[n]; ++i)
                if (coord <= dualL[n][i]) return i;
        } else {
            for (unsigned i = 1; i < N[n]; ++i)
                if (coord <= l[i]) return i - 1;
        }
        return 0;
    }

    int snap_box(const double* start, const double* stop, bool dual, unsigned* us, unsigned* ue, bool* sin, bool* ein) const
    {
        // SnapBox2Mesh, SnapMethod 0 (operator.cpp:303-327)
        double ls[3], le[3];
        for (int n = 0; n < 3; ++n) {
            ls[n] = std::fmin(start[n], stop[n]);
            le[n] = std::fmax(start[n], stop[n]);
            double mn = lines[n][0], mx = lines[n][N[n] - 1];
            if ((ls[n] < mn && le[n] < mn) || (ls[n] > mx && le[n] > mx)) return -2;
        }
        int dim = 0;
        for (int n = 0; n < 3; ++n) {
            bool a, b;
            us[n] = snap_line(n, ls[n], a, dual);
            ue[n] = snap_line(n, le[n], b, dual);
            if (sin) sin[n] = a;
            if (ein) ein[n] = b;
            if (ue[n] > us[n]) ++dim;
        }
        return dim;
    }

    int build(unsigned nrts);
    void prepare_sampling();
    double matval(int r, int kind, int ny, const double* pt) const;
    int cell_of(int n, double c) const;
    double grid_eps(const MatBox& b, const double* pt) const
    {
        const MatGrid& G = grids[b.grid];
        size_t q = 0, stride = 1;
        for (int a = 0; a < 3; ++a) {
            int k = cell_of(a, pt[a]);
            k = std::min(std::max(k, (int)G.i0[a]), (int)(G.i0[a] + G.n[a] - 1));
            q += (size_t)(k - (int)G.i0[a]) * stride;
            stride *= G.n[a];
        }
        return G.eps[q];
    }
    void effmat_at(int ny, const unsigned* p, double* E) const;
    int calc_ec();
    double timestep_var3() const;
    double timestep_var1() const;
    void ec_to_coeff(int n, size_t i);
    void apply_bc();
    void calc_pec();
    bool find_path(const double* start, const double* stop, std::vector<PathEdge>& path) const;
    void calc_pec_curves();
    int lumped_rc();
    int build_excitation();
    void build_excitation_boxes();
    void build_mur();
    void build_upml();
    void build_lorentz();
    void build_rlc();
    int build_signal(unsigned nrts);
    void grading(int n, int ny, unsigned p, const int* bcs, const unsigned* size, double& kv, double& ki) const;
};

// ---------------------------------------------------------------------------
// effective materials and EC (AverageMatQuarterCell / AverageMatCellCenter / Calc_ECPos)
// ---------------------------------------------------------------------------
void emsop::prepare_sampling()
{
    matrk.init(mats, [](const MatBox& b) { return !b.metal; });
    for (int ny = 0; ny < 3; ++ny) {
        const int n = ny, nP = (ny + 1) % 3, nPP = (ny + 2) % 3;
        Samp& S = samp[ny];
        std::vector<double> c_eps_n(N[n]), c_mu_dn(N[n]), c_mu_up(N[n]);
        std::vector<double> cPu(N[nP]), cPd(N[nP]), cPh(N[nP]), cPPu(N[nPP]), cPPd(N[nPP]), cPPh(N[nPP]);
        for (unsigned i = 0; i < N[n]; ++i) {
            c_eps_n[i] = lines[n][i] + rawD(n, i) * 0.5;
            c_mu_dn[i] = lines[n][i] - rawD(n, (int)i - 1) * 0.25;
            c_mu_up[i] = lines[n][i] + rawD(n, i) * 0.25;
        }
        for (unsigned i = 0; i < N[nP]; ++i) {
            cPu[i] = lines[nP][i] + rawD(nP, i) * 0.25;
            cPd[i] = lines[nP][i] - rawD(nP, (int)i - 1) * 0.25;
            cPh[i] = lines[nP][i] + rawD(nP, i) * 0.5;
        }
        for (unsigned i = 0; i < N[nPP]; ++i) {
            cPPu[i] = lines[nPP][i] + rawD(nPP, i) * 0.25;
            cPPd[i] = lines[nPP][i] - rawD(nPP, (int)i - 1) * 0.25;
            cPPh[i] = lines[nPP][i] + rawD(nPP, i) * 0.5;
        }
        S.c_eps_n = c_eps_n;
        S.cPu = cPu;
        S.cPd = cPd;
        S.cPh = cPh;
        S.cPPu = cPPu;
        S.cPPd = cPPd;
        S.cPPh = cPPh;
        S.c_mu_dn = c_mu_dn;
        S.c_mu_up = c_mu_up;
        S.eps_n = matrk.masks(n, c_eps_n);
        S.P_up = matrk.masks(nP, cPu);
        S.P_dn = matrk.masks(nP, cPd);
        S.PP_up = matrk.masks(nPP, cPPu);
        S.PP_dn = matrk.masks(nPP, cPPd);
        S.mu_dn = matrk.masks(n, c_mu_dn);
        S.mu_up = matrk.masks(n, c_mu_up);
        S.P_half = matrk.masks(nP, cPh);
        S.PP_half = matrk.masks(nPP, cPPh);
        S.dual_n = matrk.masks(n, dualL[n]);
        S.dual_P = matrk.masks(nP, dualL[nP]);
        S.dual_PP = matrk.masks(nPP, dualL[nPP]);
    }
}

// lowest cell index k with lines[k] <= c <= lines[k+1]
int emsop::cell_of(int n, double c) const
{
    const auto& l = lines[n];
    const size_t j = std::lower_bound(l.begin(), l.end(), c) - l.begin();
    if (j == 0) return 0;
    if (j >= l.size()) return (int)l.size() - 2;
    return (int)j - 1;
}

double emsop::matval(int r, int kind, int ny, const double* pt) const
{
    // kind 0 eps, 1 kappa, 2 mue, 3 sigma (GetMaterial, operator.cpp:1281-1334)
    if (r < 0) return kind == 0 ? bg[0] : kind == 1 ? bg[2] : kind == 2 ? bg[1] : bg[3];
    const MatBox* b = matrk.box[r];
    if (b->grid >= 0 && kind < 2) {
        const MatGrid& G = grids[b->grid];
        size_t q = 0, stride = 1;
        for (int a = 0; a < 3; ++a) {
            int k = cell_of(a, pt[a]);
            k = std::min(std::max(k, (int)G.i0[a]), (int)(G.i0[a] + G.n[a] - 1));
            q += (size_t)(k - (int)G.i0[a]) * stride;
            stride *= G.n[a];
        }
        return kind == 0 ? G.eps[q] : G.kappa[q];
    }
    switch (kind) {
    case 0: return b->eps[ny];
    case 1: return b->kappa[ny];
    case 2: return b->mue[ny];
    default: return b->sigma[ny];
    }
}

void emsop::effmat_at(int ny, const unsigned* p, double* E) const
{
    const int n = ny, nP = (ny + 1) % 3, nPP = (ny + 2) % 3;
    const unsigned W = matrk.W;
    const Samp& S = samp[ny];
    const int in = p[n], iP = p[nP], iPP = p[nPP];
    double E0, E1, E2, E3;
    if (!cellConst) {
        // AverageMatQuarterCell (operator.cpp:1447-1564)
        const uint64_t* mn = &S.eps_n[(size_t)in * W];
        int r0 = matrk.winner(mn, &S.P_up[(size_t)iP * W], &S.PP_up[(size_t)iPP * W]);
        int r1 = matrk.winner(mn, &S.P_dn[(size_t)iP * W], &S.PP_up[(size_t)iPP * W]);
        int r2 = matrk.winner(mn, &S.P_up[(size_t)iP * W], &S.PP_dn[(size_t)iPP * W]);
        int r3 = matrk.winner(mn, &S.P_dn[(size_t)iP * W], &S.PP_dn[(size_t)iPP * W]);
        double A0 = Wp(nP, iP) * Wp(nPP, iPP);
        double A1 = Wp(nP, iP - 1) * Wp(nPP, iPP);
        double A2 = Wp(nP, iP) * Wp(nPP, iPP - 1);
        double A3 = Wp(nP, iP - 1) * Wp(nPP, iPP - 1);
        double pt[4][3];
        for (int q = 0; q < 4; ++q) pt[q][n] = S.c_eps_n[in];
        pt[0][nP] = S.cPu[iP];  pt[0][nPP] = S.cPPu[iPP];
        pt[1][nP] = S.cPd[iP];  pt[1][nPP] = S.cPPu[iPP];
        pt[2][nP] = S.cPu[iP];  pt[2][nPP] = S.cPPd[iPP];
        pt[3][nP] = S.cPd[iP];  pt[3][nPP] = S.cPPd[iPP];
        double area = 0;
        E0 = matval(r0, 0, ny, pt[0]) * A0;
        E1 = matval(r0, 1, ny, pt[0]) * A0;
        area += A0;
        E0 += matval(r1, 0, ny, pt[1]) * A1;
        E1 += matval(r1, 1, ny, pt[1]) * A1;
        area += A1;
        E0 += matval(r2, 0, ny, pt[2]) * A2;
        E1 += matval(r2, 1, ny, pt[2]) * A2;
        area += A2;
        E0 += matval(r3, 0, ny, pt[3]) * A3;
        E1 += matval(r3, 1, ny, pt[3]) * A3;
        area += A3;
        E0 *= kEPS0 / area;
        E1 /= area;
        const uint64_t* mPh = &S.P_half[(size_t)iP * W];
        const uint64_t* mPPh = &S.PP_half[(size_t)iPP * W];
        int rd = matrk.winner(&S.mu_dn[(size_t)in * W], mPh, mPPh);
        int ru = matrk.winner(&S.mu_up[(size_t)in * W], mPh, mPPh);
        double pd[3], pu[3];
        pd[n] = S.c_mu_dn[in];
        pu[n] = S.c_mu_up[in];
        pd[nP] = pu[nP] = S.cPh[iP];
        pd[nPP] = pu[nPP] = S.cPPh[iPP];
        double dny = Wp(n, in - 1);
        E2 = dny / matval(rd, 2, ny, pd);
        double sg = matval(rd, 3, ny, pd);
        E3 = sg ? dny / sg : 0;
        double length = dny;
        dny = Wp(n, in);
        E2 += dny / matval(ru, 2, ny, pu);
        sg = matval(ru, 3, ny, pu);
        if (sg)
            E3 += dny / sg;
        else
            E3 = 0;
        length += dny;
        E2 = length * kMUE0 / E2;
        if (E3) E3 = length / E3;
    } else {
        // AverageMatCellCenter (operator.cpp:1336-1445)
        int lp[3] = {(int)p[0], (int)p[1], (int)p[2]};
        double area = 0;
        E0 = E1 = E2 = E3 = 0;
        const int sh[4][2] = {{0, 0}, {-1, 0}, {0, -1}, {-1, -1}};
        for (int s = 0; s < 4; ++s) {
            lp[nP] = iP + sh[s][0];
            lp[nPP] = iPP + sh[s][1];
            if (lp[nP] < 0 || lp[nPP] < 0) continue;
            int r = matrk.winner(&S.dual_n[(size_t)in * W], &S.dual_P[(size_t)lp[nP] * W], &S.dual_PP[(size_t)lp[nPP] * W]);
            double A = Wp(nP, lp[nP]) * Wp(nPP, lp[nPP]);
            double pt[3];
            pt[n] = dualL[n][in];
            pt[nP] = dualL[nP][lp[nP]];
            pt[nPP] = dualL[nPP][lp[nPP]];
            E0 += matval(r, 0, ny, pt) * A;
            E1 += matval(r, 1, ny, pt) * A;
            area += A;
        }
        E0 *= kEPS0 / area;
        E1 /= area;
        double length = 0;
        for (int s = -1; s <= 0; ++s) {
            int ln = in + s;
            if (ln < 0) continue;
            int r = matrk.winner(&S.dual_n[(size_t)ln * W], &S.dual_P[(size_t)iP * W], &S.dual_PP[(size_t)iPP * W]);
            double dny = Wp(n, ln);
            double pt[3];
            pt[n] = dualL[n][ln];
            pt[nP] = dualL[nP][iP];
            pt[nPP] = dualL[nPP][iPP];
            E2 += dny / matval(r, 2, ny, pt);
            double sg = matval(r, 3, ny, pt);
            if (sg)
                E3 += dny / sg;
            else
                E3 = 0;
            length += dny;
        }
        E2 = length * kMUE0 / E2;
        if (E3) E3 = length / E3;
    }
    E[0] = E0;
    E[1] = E1;
    E[2] = E2;
    E[3] = E3;
}

int emsop::calc_ec()
{
    const size_t tot = size();
    for (int k = 0; k < 4; ++k) ec[k].alloc(3 * tot);
    prepare_sampling();
    std::vector<int> bad(N[0], 0);
    for (int ny = 0; ny < 3; ++ny) {
        const int nP = (ny + 1) % 3, nPP = (ny + 2) % 3;
        parallel_for(N[0], nthreads, [&](unsigned x0, unsigned x1) {
            unsigned p[3];
            double E[4];
            for (p[0] = x0; p[0] < x1; ++p[0])
                for (p[1] = 0; p[1] < N[1]; ++p[1])
                    for (p[2] = 0; p[2] < N[2]; ++p[2]) {
                        effmat_at(ny, p, E);
                        for (int k = 0; k < 4; ++k)
                            if (std::isnan(E[k]) || std::isinf(E[k])) bad[p[0]] = 1;
                        const size_t q = idx(ny, p[0], p[1], p[2]);
                        // Calc_ECPos (operator.cpp:1189-1254)
                        double delta = EL(ny, p[ny], false);
                        double area = EL(nP, p[nP], true) * EL(nPP, p[nPP], true);
                        if (delta) {
                            ec[0][q] = (float)(E[0] * area / delta);
                            ec[1][q] = (float)(E[1] * area / delta);
                        } else {
                            ec[0][q] = 0.0f;
                            ec[1][q] = 0.0f;
                        }
                        delta = EL(ny, p[ny], true);
                        area = EL(nP, p[nP], false) * EL(nPP, p[nPP], false);
                        if (delta) {
                            ec[2][q] = (float)(E[2] * area / delta);
                            ec[3][q] = (float)(E[3] * area / delta);
                        } else {
                            ec[2][q] = 0.0f;
                            ec[3][q] = 0.0f;
                        }
                    }
        });
    }
    for (int b : bad)
        if (b) return fail(EMSOP_E_NUMERIC, "an effective material parameter is not finite");
    return 0;
}

// ---------------------------------------------------------------------------
// timestep (CalcTimestep_Var3 / Var1)
// ---------------------------------------------------------------------------
double emsop::timestep_var3() const
{
    const PodVec<float>& C = ec[0];
    const PodVec<float>& L = ec[2];
    std::vector<double> best(N[2], 1e200);
    parallel_for(N[2], nthreads, [&](unsigned z0, unsigned z1) {
        for (int n = 0; n < 3; ++n) {
            const int nP = (n + 1) % 3, nPP = (n + 2) % 3;
            unsigned p[3];
            for (p[2] = z0; p[2] < z1; ++p[2])
                for (p[1] = 0; p[1] < N[1]; ++p[1])
                    for (p[0] = 0; p[0] < N[0]; ++p[0]) {
                        // AdrOp with SetReflection2Cell: an out-of-range neighbour is the cell itself
                        auto at = [&](int c, int dn, int dnP, int dnPP, int) -> size_t {
                            int q[3] = {(int)p[0], (int)p[1], (int)p[2]};
                            int off[3] = {0, 0, 0};
                            off[n] += dn;
                            off[nP] += dnP;
                            off[nPP] += dnPP;
                            for (int k = 0; k < 3; ++k) {
                                int v = q[k] + off[k];
                                if (v < 0 || v > (int)N[k] - 1) v = q[k];
                                q[k] = v;
                            }
                            return idx(c, q[0], q[1], q[2]);
                        };
                        float wa = 1 / (L[at(nPP, 0, 0, 0, 0)] * C[at(n, 0, 1, 0, 0)]) + 1 / (L[at(nPP, 0, 0, 0, 0)] * C[at(n, 0, 0, 0, 0)]);
                        double wqp = wa;
                        float wb = 1 / (L[at(nP, 0, 0, 0, 0)] * C[at(n, 0, 0, 1, 0)]) + 1 / (L[at(nP, 0, 0, 0, 0)] * C[at(n, 0, 0, 0, 0)]);
                        wqp += wb;
                        float wc = 1 / (L[at(nPP, 0, -1, 0, 0)] * C[at(n, 0, 0, 0, 0)]) + 1 / (L[at(nPP, 0, -1, 0, 0)] * C[at(n, 0, -1, 0, 0)]);
                        wqp += wc;
                        float wd = 1 / (L[at(nP, 0, -1, -1, 0)] * C[at(n, 0, -1, 0, 0)]) + 1 / (L[at(nP, 0, -1, -1, 0)] * C[at(n, 0, -1, -1, 0)]);
                        wqp += wd;
                        double w4[4];
                        w4[0] = 1 / (L[at(nPP, 0, 0, 0, 0)] * C[at(nP, 0, 0, 0, 0)]);
                        w4[1] = 1 / (L[at(nPP, 0, -1, 0, 0)] * C[at(nP, 0, 0, 0, 0)]);
                        w4[2] = 1 / (L[at(nP, 0, 0, 0, 0)] * C[at(nPP, 0, 0, 0, 0)]);
                        w4[3] = 1 / (L[at(nP, 0, 0, -1, 0)] * C[at(nPP, 0, 0, 0, 0)]);
                        double mn = w4[0];
                        for (int k = 1; k < 4; ++k)
                            if (w4[k] < mn) mn = w4[k];
                        double wt1 = w4[0] + w4[1] + w4[2] + w4[3] - 2 * mn;
                        w4[0] = 1 / (L[at(nPP, 0, 0, 0, 0)] * C[at(nP, 1, 0, 0, 0)]);
                        w4[1] = 1 / (L[at(nPP, 0, -1, 0, 0)] * C[at(nP, 1, 0, 0, 0)]);
                        w4[2] = 1 / (L[at(nP, 0, 0, 0, 0)] * C[at(nPP, 1, 0, 0, 0)]);
                        w4[3] = 1 / (L[at(nP, 0, 0, -1, 0)] * C[at(nPP, 1, 0, 0, 0)]);
                        mn = w4[0];
                        for (int k = 1; k < 4; ++k)
                            if (w4[k] < mn) mn = w4[k];
                        double wt2 = w4[0] + w4[1] + w4[2] + w4[3] - 2 * mn;
                        double w_total = wqp + wt1 + wt2;
                        double newT = 2 / std::sqrt(w_total);
                        if ((newT < best[p[2]]) && (newT > 0.0)) best[p[2]] = newT;
                    }
        }
    });
    double dt = 1e200;
    for (double b : best)
        if (b < dt) dt = b;
    return dt;
}

double emsop::timestep_var1() const
{
    const PodVec<float>& C = ec[0];
    const PodVec<float>& L = ec[2];
    double dt = 1e200;
    unsigned p[3];
    for (int n = 0; n < 3; ++n) {
        const int nP = (n + 1) % 3, nPP = (n + 2) % 3;
        for (p[2] = 0; p[2] < N[2]; ++p[2])
            for (p[1] = 0; p[1] < N[1]; ++p[1])
                for (p[0] = 0; p[0] < N[0]; ++p[0]) {
                    unsigned qPM[3] = {p[0], p[1], p[2]}, qPPM[3] = {p[0], p[1], p[2]};
                    if (qPM[nP] > 0) --qPM[nP];
                    if (qPPM[nPP] > 0) --qPPM[nPP];
                    float s = (4 / L[idx(nP, p[0], p[1], p[2])] + 4 / L[idx(nP, qPPM[0], qPPM[1], qPPM[2])] +
                               4 / L[idx(nPP, p[0], p[1], p[2])] + 4 / L[idx(nPP, qPM[0], qPM[1], qPM[2])]) /
                              C[idx(n, p[0], p[1], p[2])];
                    double newT = 2 / std::sqrt((double)s);
                    if ((newT < dt) && (newT > 0.0)) dt = newT;
                }
    }
    return dt;
}

// Calc_ECOperatorPos (operator.cpp:956-984)
void emsop::ec_to_coeff(int n, size_t i)
{
    (void)n;
    double C = ec[0][i], G = ec[1][i];
    if (C > 0) {
        coeff[0][i] = (float)((1.0 - dT * G / 2.0 / C) / (1.0 + dT * G / 2.0 / C));
        coeff[1][i] = (float)((dT / C) / (1.0 + dT * G / 2.0 / C));
    } else {
        coeff[0][i] = 0;
        coeff[1][i] = 0;
    }
    double L = ec[2][i], R = ec[3][i];
    if (L > 0) {
        coeff[2][i] = (float)((1.0 - dT * R / 2.0 / L) / (1.0 + dT * R / 2.0 / L));
        coeff[3][i] = (float)((dT / L) / (1.0 + dT * R / 2.0 / L));
    } else {
        coeff[2][i] = 0;
        coeff[3][i] = 0;
    }
}

void emsop::apply_bc()
{
    // ApplyElectricBC with PEC except BC == -1 (operator.cpp:1042-1048, 1099-1137)
    auto& vv = coeff[0];
    auto& vi = coeff[1];
    for (int n = 0; n < 3; ++n) {
        const int nP = (n + 1) % 3, nPP = (n + 2) % 3;
        unsigned p[3];
        for (p[nP] = 0; p[nP] < N[nP]; ++p[nP])
            for (p[nPP] = 0; p[nPP] < N[nPP]; ++p[nPP]) {
                if (bc[2 * n] != -1) {
                    p[n] = 0;
                    for (int c : {nP, nPP}) {
                        vv[idx(c, p[0], p[1], p[2])] = 0;
                        vi[idx(c, p[0], p[1], p[2])] = 0;
                    }
                }
                if (bc[2 * n + 1] != -1) {
                    p[n] = N[n] - 1;
                    for (int c : {n, nP, nPP}) {
                        vv[idx(c, p[0], p[1], p[2])] = 0;
                        vi[idx(c, p[0], p[1], p[2])] = 0;
                    }
                }
            }
    }
}

void emsop::calc_pec()
{
    // CalcPEC_Range (operator.cpp:2045-2083): MATERIAL|METAL lookup at the Yee edge centre
    bool any_metal = false;
    for (const auto& m : mats) any_metal |= m.metal;
    if (!any_metal) return;
    Ranked<MatBox> rk;
    rk.init(mats, [](const MatBox&) { return true; });
    const unsigned W = rk.W;
    for (int n = 0; n < 3; ++n) {
        std::vector<uint64_t> m[3];
        const std::vector<double>* c[3];
        for (int a = 0; a < 3; ++a) {
            c[a] = a == n ? &dualL[a] : &lines[a];
            m[a] = rk.masks(a, *c[a]);
        }
        for (unsigned x = 0; x < N[0]; ++x)
            for (unsigned y = 0; y < N[1]; ++y)
                for (unsigned z = 0; z < N[2]; ++z) {
                    const double P[3] = {(*c[0])[x], (*c[1])[y], (*c[2])[z]};
                    int r = rk.winner_at(&m[0][(size_t)x * W], &m[1][(size_t)y * W], &m[2][(size_t)z * W], P);
                    if (r >= 0 && rk.box[r]->metal) {
                        coeff[0][idx(n, x, y, z)] = 0;
                        coeff[1][idx(n, x, y, z)] = 0;
                    }
                }
    }
}


// Operator::FindPath (operator.cpp:390-496) with SnapLine2Mesh
// (operator.cpp:360-388): a staircase of mesh edges from the node nearest
// start to the node nearest stop.  Each step moves to the neighbouring node
// whose foot on the line lies further along and whose distance from the line
// is smallest (the first candidate wins a tie, axes x, y, z, down before up);
// then at most one edge per axis closes the gap to the stop node.  A path
// that leaves the mesh is dropped (empty), as is a line wholly beyond one end
// of an axis.
bool emsop::find_path(const double* start, const double* stop, std::vector<PathEdge>& path) const
{
    path.clear();
    unsigned us[3], ue[3];
    for (int n = 0; n < 3; ++n) {
        bool in;
        us[n] = snap_line(n, start[n], in, false);
        ue[n] = snap_line(n, stop[n], in, false);
    }
    for (int n = 0; n < 3; ++n) {
        const double lo = lines[n][0], hi = lines[n][N[n] - 1];
        if ((start[n] < lo && stop[n] < lo) || (start[n] > hi && stop[n] > hi)) return false;
    }
    unsigned cur[3] = {us[0], us[1], us[2]};
    const double ms[3] = {lines[0][us[0]], lines[1][us[1]], lines[2][us[2]]};
    const double me[3] = {lines[0][ue[0]], lines[1][ue[1]], lines[2][ue[2]]};
    double foot = 0, dist = 0, startFoot, stopFoot;
    point_line_distance(ms, start, stop, startFoot, dist);
    point_line_distance(me, start, stop, stopFoot, dist);
    double currFoot = startFoot, minFoot = startFoot, minDist = 0;
    int minDir = 0;
    bool up = false;
    while (minFoot < stopFoot) {
        minDist = 1e300;
        for (int n = 0; n < 3; ++n) {
            double P[3] = {lines[0][cur[0]], lines[1][cur[1]], lines[2][cur[2]]};
            if (cur[n] >= 1) {
                P[n] = lines[n][cur[n] - 1];
                point_line_distance(P, start, stop, foot, dist);
                if (foot > currFoot && dist < minDist) {
                    minFoot = foot;
                    minDist = dist;
                    minDir = n;
                    up = false;
                }
            }
            if (cur[n] + 1 < N[n]) {
                P[n] = lines[n][cur[n] + 1];
                point_line_distance(P, start, stop, foot, dist);
                if (foot > currFoot && dist < minDist) {
                    minFoot = foot;
                    minDist = dist;
                    minDir = n;
                    up = true;
                }
            }
        }
        PathEdge e{{cur[0], cur[1], cur[2]}, minDir};
        if (up) {
            cur[minDir] += 1;
        } else {
            cur[minDir] -= 1;   // unsigned: below line 0 wraps and fails the check below
            e.pos[minDir] -= 1;
        }
        for (int n = 0; n < 3; ++n)
            if (cur[n] >= N[n]) {
                path.clear();
                return false;
            }
        path.push_back(e);
        currFoot = minFoot;
    }
    for (int n = 0; n < 3; ++n) {
        if (cur[n] > ue[n]) {
            --cur[n];
            path.push_back(PathEdge{{cur[0], cur[1], cur[2]}, n});
        } else if (cur[n] < ue[n]) {
            path.push_back(PathEdge{{cur[0], cur[1], cur[2]}, n});
        }
    }
    return !path.empty();
}

void emsop::calc_pec_curves()
{
    // CalcPEC_Curves (operator.cpp:2086-2118): every segment of a metal curve
    std::vector<PathEdge> path;
    for (const auto& c : curves) {
        if (c.exc) continue;
        for (size_t i = 1; i < c.pts.size() / 3; ++i) {
            find_path(&c.pts[3 * (i - 1)], &c.pts[3 * i], path);
            for (const auto& e : path) {
                const size_t q = idx(e.dir, e.pos[0], e.pos[1], e.pos[2]);
                coeff[0][q] = 0;
                coeff[1][q] = 0;
            }
        }
    }
}

int emsop::lumped_rc()
{
    // Calc_LumpedElements (operator.cpp:1586-1763) for parallel-RC elements
    for (const auto& le : lumped) {
        if (!le.parallel || le.L > 0) continue;
        double Cap = le.C, R = le.R;
        if (Cap <= 0) Cap = NAN;
        if (R < 0) R = NAN;
        if (std::isnan(R) && std::isnan(Cap)) continue;
        const int ny = le.dir;
        if (ny < 0 || ny > 2) continue;
        const int nyP = (ny + 1) % 3, nyPP = (ny + 2) % 3;
        unsigned s[3], e[3];
        int dim = snap_box(le.lo, le.hi, false, s, e, nullptr, nullptr);
        if (dim <= 0) continue;
        if (s[ny] == e[ny]) continue;
        unsigned p[3];
        double unitGC = 0;
        for (p[ny] = s[ny]; p[ny] < e[ny]; ++p[ny]) {
            double plane = 0;
            for (p[nyP] = s[nyP]; p[nyP] <= e[nyP]; ++p[nyP])
                for (p[nyPP] = s[nyPP]; p[nyPP] <= e[nyPP]; ++p[nyPP])
                    plane += EL(nyP, p[nyP], true) * EL(nyPP, p[nyPP], true) / EL(ny, p[ny], false);
            unitGC += 1 / plane;
        }
        unitGC = 1 / unitGC;
        double kappa = 0, eps = 0;
        if (R > 0) kappa = 1 / R / unitGC;
        if (Cap > 0) {
            eps = Cap / unitGC;
            if (eps < kEPS0) Cap = 0;
        }
        for (p[ny] = s[ny]; p[ny] < e[ny]; ++p[ny])
            for (p[nyP] = s[nyP]; p[nyP] <= e[nyP]; ++p[nyP])
                for (p[nyPP] = s[nyPP]; p[nyPP] <= e[nyPP]; ++p[nyPP]) {
                    size_t q = idx(ny, p[0], p[1], p[2]);
                    double A = EL(nyP, p[nyP], true) * EL(nyPP, p[nyPP], true);
                    if (Cap > 0) ec[0][q] = (float)(eps * A / EL(ny, p[ny], false));
                    if (R > 0) ec[1][q] = (float)(kappa * A / EL(ny, p[ny], false));
                    if (R == 0) {
                        coeff[0][q] = 0;
                        coeff[1][q] = 0;
                    } else
                        ec_to_coeff(ny, q);
                }
        if (le.caps) {
            for (p[nyP] = s[nyP]; p[nyP] <= e[nyP]; ++p[nyP])
                for (p[nyPP] = s[nyPP]; p[nyPP] <= e[nyPP]; ++p[nyPP])
                    for (unsigned pl : {s[ny], e[ny]}) {
                        p[ny] = pl;
                        if (p[nyP] < e[nyP]) {
                            coeff[0][idx(nyP, p[0], p[1], p[2])] = 0;
                            coeff[1][idx(nyP, p[0], p[1], p[2])] = 0;
                        }
                        if (p[nyPP] < e[nyPP]) {
                            coeff[0][idx(nyPP, p[0], p[1], p[2])] = 0;
                            coeff[1][idx(nyPP, p[0], p[1], p[2])] = 0;
                        }
                    }
        }
    }
    return 0;
}

void emsop_apply_magnetic_bc(emsop* h)
{
    // ApplyMagneticBC (operator.cpp:1139-1187): PMC where BC == 1, last current line always zero
    auto& ii = h->coeff[2];
    auto& iv = h->coeff[3];
    const unsigned* N = h->N;
    for (int n = 0; n < 3; ++n) {
        const int nP = (n + 1) % 3, nPP = (n + 2) % 3;
        unsigned p[3];
        for (p[nP] = 0; p[nP] < N[nP]; ++p[nP])
            for (p[nPP] = 0; p[nPP] < N[nPP]; ++p[nPP]) {
                if (h->bc[2 * n] == 1) {
                    p[n] = 0;
                    for (int c : {n, nP, nPP}) {
                        ii[h->idx(c, p[0], p[1], p[2])] = 0;
                        iv[h->idx(c, p[0], p[1], p[2])] = 0;
                    }
                }
                if (h->bc[2 * n + 1] == 1) {
                    p[n] = N[n] - 2;
                    for (int c : {nP, nPP}) {
                        ii[h->idx(c, p[0], p[1], p[2])] = 0;
                        iv[h->idx(c, p[0], p[1], p[2])] = 0;
                    }
                }
                p[n] = N[n] - 1;
                for (int c : {n, nP, nPP}) {
                    ii[h->idx(c, p[0], p[1], p[2])] = 0;
                    iv[h->idx(c, p[0], p[1], p[2])] = 0;
                }
            }
    }
}

// ---------------------------------------------------------------------------
// excitation extension (operator_ext_excitation.cpp:105-297), box primitives
// ---------------------------------------------------------------------------
int emsop::build_excitation()
{
    exc[0] = ExcList();
    exc[1] = ExcList();
    if (!excs.empty()) build_excitation_boxes();
    // curves after the boxes (operator_ext_excitation.cpp:235-287): voltage
    // excitation on the path edges, type 0 (soft) or 1 (hard, zeroes vv / vi)
    std::vector<PathEdge> path;
    for (const auto& c : curves) {
        if (!c.exc) continue;
        for (size_t i = 1; i < c.pts.size() / 3; ++i) {
            find_path(&c.pts[3 * (i - 1)], &c.pts[3 * i], path);
            for (const auto& e : path) {
                const int n = e.dir;
                if (e.pos[n] >= N[n] - 1 || (c.type != 0 && c.type != 1)) continue;
                double P[3];
                for (int a = 0; a < 3; ++a) P[a] = (a == n ? dualL[a] : lines[a])[e.pos[a]];
                const double amp = weighted(c.weight, n, c.val[n], P) * EL(n, e.pos[n], false);
                if (amp != 0) {
                    for (int a = 0; a < 3; ++a) exc[0].idx[a].push_back(e.pos[a]);
                    exc[0].dir.push_back((unsigned short)n);
                    exc[0].amp.push_back((float)amp);
                    exc[0].delay.push_back((unsigned)(c.delay / dT));
                }
                if (c.type == 1) {
                    const size_t q = idx(n, e.pos[0], e.pos[1], e.pos[2]);
                    coeff[0][q] = 0;
                    coeff[1][q] = 0;
                }
            }
        }
    }
    return 0;
}

void emsop::build_excitation_boxes()
{
    Ranked<ExcBox> rk;
    rk.init(excs, [](const ExcBox&) { return true; });
    const unsigned W = rk.W;
    std::vector<uint64_t> mP[3], mD[3];
    for (int a = 0; a < 3; ++a) {
        mP[a] = rk.masks(a, lines[a]);
        mD[a] = rk.masks(a, dualL[a]);
    }
    auto anyset = [&](const std::vector<uint64_t>& m, unsigned i) {
        for (unsigned w = 0; w < W; ++w)
            if (m[(size_t)i * W + w]) return true;
        return false;
    };
    auto& vv = coeff[0];
    auto& vi = coeff[1];
    auto& ii = coeff[2];
    auto& iv = coeff[3];
    unsigned p[3];
    for (p[2] = 0; p[2] < N[2]; ++p[2]) {
        if (!anyset(mP[2], p[2]) && !anyset(mD[2], p[2])) continue;
        for (p[1] = 0; p[1] < N[1]; ++p[1]) {
            if (!anyset(mP[1], p[1]) && !anyset(mD[1], p[1])) continue;
            for (p[0] = 0; p[0] < N[0]; ++p[0]) {
                if (!anyset(mP[0], p[0]) && !anyset(mD[0], p[0])) continue;
                // electric field excite
                for (int n = 0; n < 3; ++n) {
                    if (p[n] >= N[n] - 1) continue;  // GetYeeCoords(...,false) == false
                    const std::vector<uint64_t>* m[3];
                    double P[3];
                    for (int a = 0; a < 3; ++a) {
                        m[a] = (a == n) ? &mD[a] : &mP[a];
                        P[a] = ((a == n) ? dualL[a] : lines[a])[p[a]];
                    }
                    int r = rk.winner_at(&(*m[0])[(size_t)p[0] * W], &(*m[1])[(size_t)p[1] * W],
                                         &(*m[2])[(size_t)p[2] * W], P);
                    if (r < 0) continue;
                    const ExcBox* b = rk.box[r];
                    if (b->type != 0 && b->type != 1) continue;
                    double amp = weighted(b->weight, n, b->val[n], P) * EL(n, p[n], false);
                    if (amp != 0) {
                        for (int a = 0; a < 3; ++a) exc[0].idx[a].push_back(p[a]);
                        exc[0].dir.push_back((unsigned short)n);
                        exc[0].amp.push_back((float)amp);
                        exc[0].delay.push_back((unsigned)(b->delay / dT));
                    }
                    if (b->type == 1) {
                        vv[idx(n, p[0], p[1], p[2])] = 0;
                        vi[idx(n, p[0], p[1], p[2])] = 0;
                    }
                }
                // magnetic field excite
                for (int n = 0; n < 3; ++n) {
                    if (p[0] >= N[0] - 1 || p[1] >= N[1] - 1 || p[2] >= N[2] - 1) continue;
                    const std::vector<uint64_t>* m[3];
                    double P[3];
                    for (int a = 0; a < 3; ++a) {
                        m[a] = (a == n) ? &mP[a] : &mD[a];
                        P[a] = ((a == n) ? lines[a] : dualL[a])[p[a]];
                    }
                    int r = rk.winner_at(&(*m[0])[(size_t)p[0] * W], &(*m[1])[(size_t)p[1] * W],
                                         &(*m[2])[(size_t)p[2] * W], P);
                    if (r < 0) continue;
                    const ExcBox* b = rk.box[r];
                    if (b->type != 2 && b->type != 3) continue;
                    double amp = weighted(b->weight, n, b->val[n], P) * EL(n, p[n], true);
                    if (amp != 0) {
                        for (int a = 0; a < 3; ++a) exc[1].idx[a].push_back(p[a]);
                        exc[1].dir.push_back((unsigned short)n);
                        exc[1].amp.push_back((float)amp);
                        exc[1].delay.push_back((unsigned)(b->delay / dT));
                    }
                    if (b->type == 3) {
                        ii[idx(n, p[0], p[1], p[2])] = 0;
                        iv[idx(n, p[0], p[1], p[2])] = 0;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Mur ABC (operator_ext_mur_abc.cpp:80-186)
// ---------------------------------------------------------------------------
void emsop::build_mur()
{
    mur.clear();
    Ranked<MatBox> rk;
    rk.init(mats, [](const MatBox& b) { return !b.metal; });
    const unsigned W = rk.W;
    for (int f = 0; f < 6; ++f) {
        if (bc[f] != 2) continue;
        MurFace F;
        const int ny = f / 2;
        const bool top = f % 2;
        const int nyP = (ny + 1) % 3, nyPP = (ny + 2) % 3;
        F.ny = ny;
        F.line = top ? N[ny] - 1 : 0;
        F.shift = top ? N[ny] - 2 : 1;
        double delta = std::fabs(EL(ny, F.line, false));
        std::vector<double> cn(1);
        if (F.line == 0)
            cn[0] = lines[ny][F.line] + delta / 2 / gd;
        else
            cn[0] = lines[ny][F.line] - delta / 2 / gd;
        auto mN = rk.masks(ny, cn);
        auto mA = rk.masks(nyP, lines[nyP]);
        auto mB = rk.masks(nyPP, lines[nyPP]);
        const unsigned na = N[nyP], nb = N[nyPP];
        F.cP.assign((size_t)na * nb, 0);
        F.cPP.assign((size_t)na * nb, 0);
        for (unsigned i = 0; i < na; ++i)
            for (unsigned j = 0; j < nb; ++j) {
                const uint64_t* m3[3];
                m3[ny] = &mN[0];
                m3[nyP] = &mA[(size_t)i * W];
                m3[nyPP] = &mB[(size_t)j * W];
                int r = rk.winner(m3[0], m3[1], m3[2]);
                double c0t;
                size_t k = (size_t)i * nb + j;
                if (r >= 0) {
                    const MatBox* b = rk.box[r];
                    double pt[3];
                    pt[ny] = cn[0];
                    pt[nyP] = lines[nyP][i];
                    pt[nyPP] = lines[nyPP][j];
                    const double geps = b->grid >= 0 ? grid_eps(*b, pt) : 0.0;
                    double eps = b->grid >= 0 ? geps : b->eps[nyP], mue = b->mue[nyP];
                    if (vph[f] > 0.0)
                        c0t = vph[f] * dT;
                    else
                        c0t = kC0 * dT / std::sqrt(eps * mue);
                    F.cP[k] = (float)((c0t - delta) / (c0t + delta));
                    eps = b->grid >= 0 ? geps : b->eps[nyPP];
                    mue = b->mue[nyPP];
                    if (vph[f] > 0.0)
                        c0t = vph[f] * dT;
                    else
                        c0t = kC0 * dT / std::sqrt(eps * mue);
                    F.cPP[k] = (float)((c0t - delta) / (c0t + delta));
                } else {
                    if (vph[f] > 0.0)
                        c0t = vph[f] * dT;
                    else
                        c0t = kC0 / std::sqrt(bg[0] * bg[1]) * dT;
                    F.cP[k] = (float)((c0t - delta) / (c0t + delta));
                    F.cPP[k] = F.cP[k];
                }
            }
        // Engine_Ext_Mur_ABC: start late when a voltage excitation lies on the face
        int maxDelay = -1;
        for (size_t q = 0; q < exc[0].dir.size(); ++q)
            if ((exc[0].dir[q] == nyP || exc[0].dir[q] == nyPP) && exc[0].idx[ny][q] == F.line)
                if ((int)exc[0].delay[q] > maxDelay) maxDelay = (int)exc[0].delay[q];
        F.start_ts = (maxDelay >= 0) ? (unsigned)(maxDelay + (int)sigLen + 10) : 0;
        mur.push_back(std::move(F));
    }
}

// ---------------------------------------------------------------------------
// UPML (operator_ext_upml.cpp:69-445)
// ---------------------------------------------------------------------------
static double upml_grading(double D, double dl, double Wd, double Z)
{
    // default grading function (operator_ext_upml.cpp:30) written out left to right
    return -std::log(1e-6) * std::log(2.5) / (2 * dl * Z * (std::pow(2.5, Wd / dl) - 1)) * std::pow(2.5, D / dl);
}

void emsop::grading(int n, int ny, unsigned p, const int* bcs, const unsigned* size, double& kv, double& ki) const
{
    // one axis of CalcGradingKappa (operator_ext_upml.cpp:269-343)
    kv = ki = 0;
    const auto& L = lines[n];
    if (p <= size[2 * n] && bcs[2 * n] == 3) {
        double width = (L[size[2 * n]] - L[0]) * gd;
        double depth = width - (L[p] - L[0]) * gd;
        double el = EL(n, p, false);
        if (n == ny) depth -= el / 2;
        double dl = width / size[2 * n];
        kv = depth > 0 ? upml_grading(depth, dl, width, kZ0) : 0;
        if (n == ny) depth += el / 2;
        if (n != ny) depth -= el / 2;
        if (depth < 0) depth = 0;
        ki = depth > 0 ? upml_grading(depth, dl, width, kZ0) : 0;
    } else if (p >= N[n] - 1 - size[2 * n + 1] && bcs[2 * n + 1] == 3) {
        double width = (L[N[n] - 1] - L[N[n] - size[2 * n + 1] - 1]) * gd;
        double depth = width - (L[N[n] - 1] - L[p]) * gd;
        double el = EL(n, p, false);
        if (n == ny) depth += el / 2;
        double dl = width / (size[2 * n]);  // the reference divides by the lower size here too
        kv = depth > 0 ? upml_grading(depth, dl, width, kZ0) : 0;
        if (n == ny) depth -= el / 2;
        if (n != ny) depth += el / 2;
        if (depth > width) depth = 0;
        ki = depth > 0 ? upml_grading(depth, dl, width, kZ0) : 0;
    }
}

void emsop::build_upml()
{
    upml.clear();
    int B[6];
    unsigned S[6];
    for (int k = 0; k < 6; ++k) {
        B[k] = bc[k];
        S[k] = pml[k];
    }
    for (int n = 0; n < 3; ++n)
        if (S[2 * n] * (B[2 * n] == 3) + S[2 * n + 1] * (B[2 * n + 1] == 3) >= N[n]) {
            B[2 * n] = B[2 * n + 1] = 0;
            S[2 * n] = S[2 * n + 1] = 0;
        }
    std::vector<std::pair<std::array<unsigned, 3>, std::array<unsigned, 3>>> ranges;
    unsigned st[3] = {0, 0, 0}, sp[3] = {N[0] - 1, N[1] - 1, N[2] - 1};
    auto push = [&]() { ranges.push_back({{st[0], st[1], st[2]}, {sp[0], sp[1], sp[2]}}); };
    if (B[0] == 3) { st[0] = 0; sp[0] = S[0]; push(); }
    if (B[1] == 3) { st[0] = N[0] - 1 - S[1]; sp[0] = N[0] - 1; push(); }
    st[0] = (S[0] + 1) * (B[0] == 3);
    sp[0] = N[0] - 1 - (S[0] + 1) * (B[1] == 3);
    if (B[2] == 3) { st[1] = 0; sp[1] = S[2]; push(); }
    if (B[3] == 3) { st[1] = N[1] - 1 - S[3]; sp[1] = N[1] - 1; push(); }
    st[1] = (S[2] + 1) * (B[2] == 3);
    sp[1] = N[1] - 1 - (S[3] + 1) * (B[3] == 3);
    if (B[4] == 3) { st[2] = 0; sp[2] = S[4]; push(); }
    if (B[5] == 3) { st[2] = N[2] - 1 - S[5]; sp[2] = N[2] - 1; push(); }

    const double E2 = 2 * kEPS0;
    auto& vv = coeff[0];
    auto& vi = coeff[1];
    auto& ii = coeff[2];
    auto& iv = coeff[3];
    for (auto& r : ranges) {
        Slab s;
        for (int k = 0; k < 3; ++k) {
            s.start[k] = r.first[k];
            s.num[k] = r.second[k] - r.first[k] + 1;
        }
        const size_t tot = (size_t)3 * s.num[0] * s.num[1] * s.num[2];
        for (int k = 0; k < 6; ++k) s.a[k].assign(tot, 0.0f);
        // per-axis grading tables: g[ny][axis][local] = (kv, ki)
        std::vector<double> gv[3][3], gi[3][3];
        for (int ny = 0; ny < 3; ++ny)
            for (int a = 0; a < 3; ++a) {
                gv[ny][a].resize(s.num[a]);
                gi[ny][a].resize(s.num[a]);
                for (unsigned l = 0; l < s.num[a]; ++l) grading(a, ny, l + s.start[a], B, S, gv[ny][a][l], gi[ny][a][l]);
            }
        auto li = [&](int n, unsigned x, unsigned y, unsigned z) {
            return (((size_t)n * s.num[0] + x) * s.num[1] + y) * s.num[2] + z;
        };
        parallel_for(s.num[0], nthreads, [&](unsigned x0, unsigned x1) {
            unsigned loc[3], p[3];
            for (loc[0] = x0; loc[0] < x1; ++loc[0])
                for (loc[1] = 0; loc[1] < s.num[1]; ++loc[1])
                    for (loc[2] = 0; loc[2] < s.num[2]; ++loc[2]) {
                        for (int k = 0; k < 3; ++k) p[k] = loc[k] + s.start[k];
                        for (int n = 0; n < 3; ++n) {
                            const int nP = (n + 1) % 3, nPP = (n + 2) % 3;
                            double E[4];
                            effmat_at(n, p, E);
                            double kv[3], ki[3];
                            for (int a = 0; a < 3; ++a) {
                                kv[a] = gv[n][a][loc[a]];
                                ki[a] = gi[n][a][loc[a]];
                            }
                            const size_t q = idx(n, p[0], p[1], p[2]);
                            const size_t l = li(n, loc[0], loc[1], loc[2]);
                            if (((kv[0] + kv[1] + kv[2]) != 0) && (E[1] < 1e3)) {
                                if ((vv[q] + vi[q]) != 0) {
                                    vv[q] = (float)((E2 - kv[nP] * dT) / (E2 + kv[nP] * dT));
                                    vi[q] = (float)((E2 * dT) / (E2 + kv[nP] * dT) * EL(n, p[n], false) /
                                                    (EL(nP, p[nP], true) * EL(nPP, p[nPP], true)));
                                    s.a[0][l] = (float)((E2 - kv[nPP] * dT) / (E2 + kv[nPP] * dT));
                                    s.a[2][l] = (float)((E2 + kv[n] * dT) / (E2 + kv[nPP] * dT) / E[0]);
                                    s.a[1][l] = (float)((E2 - kv[n] * dT) / (E2 + kv[nPP] * dT) / E[0]);
                                }
                            } else {
                                s.a[0][l] = vv[q];
                                vv[q] = 0;
                                s.a[1][l] = 0;
                                s.a[2][l] = 1;
                            }
                            if ((ki[0] + ki[1] + ki[2]) != 0) {
                                if ((ii[q] + iv[q]) != 0) {
                                    ii[q] = (float)((E2 - ki[nP] * dT) / (E2 + ki[nP] * dT));
                                    iv[q] = (float)((E2 * dT) / (E2 + ki[nP] * dT) * EL(n, p[n], true) /
                                                    (EL(nP, p[nP], false) * EL(nPP, p[nPP], false)));
                                    s.a[3][l] = (float)((E2 - ki[nPP] * dT) / (E2 + ki[nPP] * dT));
                                    s.a[5][l] = (float)((E2 + ki[n] * dT) / (E2 + ki[nPP] * dT) / E[2]);
                                    s.a[4][l] = (float)((E2 - ki[n] * dT) / (E2 + ki[nPP] * dT) / E[2]);
                                }
                            } else {
                                s.a[3][l] = ii[q];
                                ii[q] = 0;
                                s.a[4][l] = 0;
                                s.a[5][l] = 1;
                            }
                        }
                    }
        });
        upml.push_back(std::move(s));
    }
}

// ---------------------------------------------------------------------------
// excitation signal (FDTD/excitation.cpp:150-276)
// ---------------------------------------------------------------------------
// Expression of t for custom excitation signals, in the syntax of the
// function parser the reference uses (fparser; Excitation::CalcCustomExcitation,
// FDTD/excitation.cpp:219-252): numbers, t, the constants pi and e, + - * / ^
// (power, right-associative, binding tighter than a leading minus),
// comparisons (< <= > >= = !=, giving 1 or 0), & and | (logical), !, the
// functions sin cos tan asin acos atan sinh cosh tanh exp log log10 sqrt
// abs floor ceil int (round to nearest), atan2 pow min max (two arguments)
// and if(c, a, b).  Parsed once into a tree and evaluated in double, left to
// right as written.
namespace {
struct Expr {
    enum Op { NUM, VAR, NEG, NOT, ADD, SUB, MUL, DIV, POW, LT, LE, GT, GE, EQ, NE, AND, OR, FN1, FN2, IF };
    Op op;
    double v = 0;
    int fn = 0;
    std::vector<std::unique_ptr<Expr>> a;
};

struct ExprParser {
    const std::string& s;
    const std::vector<std::string>& vars;   // variable names; VAR nodes hold the index in fn
    size_t i = 0;
    std::string err;
    ExprParser(const std::string& str, const std::vector<std::string>& v) : s(str), vars(v) {}
    void ws() { while (i < s.size() && std::isspace((unsigned char)s[i])) ++i; }
    bool eat(const char* tok)
    {
        ws();
        size_t n = std::strlen(tok);
        if (s.compare(i, n, tok) == 0) {
            i += n;
            return true;
        }
        return false;
    }
    std::unique_ptr<Expr> mk(Expr::Op op, std::unique_ptr<Expr> x, std::unique_ptr<Expr> y = nullptr)
    {
        auto e = std::make_unique<Expr>();
        e->op = op;
        e->a.push_back(std::move(x));
        if (y) e->a.push_back(std::move(y));
        return e;
    }
    std::unique_ptr<Expr> fail_(const std::string& m)
    {
        if (err.empty()) err = m + " at position " + std::to_string(i);
        return nullptr;
    }
    // or := and ('|' and)* ; and := cmp ('&' cmp)* ; cmp := sum (relop sum)?
    std::unique_ptr<Expr> orx()
    {
        auto x = andx();
        while (x && eat("|")) x = mk(Expr::OR, std::move(x), andx());
        return x;
    }
    std::unique_ptr<Expr> andx()
    {
        auto x = cmp();
        while (x && eat("&")) x = mk(Expr::AND, std::move(x), cmp());
        return x;
    }
    std::unique_ptr<Expr> cmp()
    {
        auto x = sum();
        if (!x) return x;
        static const std::pair<const char*, Expr::Op> ops[] = {{"<=", Expr::LE}, {">=", Expr::GE}, {"!=", Expr::NE},
                                                               {"<", Expr::LT},  {">", Expr::GT},  {"=", Expr::EQ}};
        for (const auto& o : ops)
            if (eat(o.first)) return mk(o.second, std::move(x), sum());
        return x;
    }
    std::unique_ptr<Expr> sum()
    {
        auto x = prod();
        while (x) {
            if (eat("+")) x = mk(Expr::ADD, std::move(x), prod());
            else if (eat("-")) x = mk(Expr::SUB, std::move(x), prod());
            else break;
        }
        return x;
    }
    std::unique_ptr<Expr> prod()
    {
        auto x = unary();
        while (x) {
            if (eat("*")) x = mk(Expr::MUL, std::move(x), unary());
            else if (eat("/")) x = mk(Expr::DIV, std::move(x), unary());
            else break;
        }
        return x;
    }
    std::unique_ptr<Expr> unary()
    {
        if (eat("-")) {
            auto x = unary();
            return x ? mk(Expr::NEG, std::move(x)) : nullptr;
        }
        if (eat("!")) {
            auto x = unary();
            return x ? mk(Expr::NOT, std::move(x)) : nullptr;
        }
        return power();
    }
    std::unique_ptr<Expr> power()
    {
        auto x = atom();
        if (x && eat("^")) {
            auto y = unary();   // right-associative; the exponent may carry a sign
            return y ? mk(Expr::POW, std::move(x), std::move(y)) : nullptr;
        }
        return x;
    }
    std::unique_ptr<Expr> atom()
    {
        ws();
        if (i >= s.size()) return fail_("unexpected end");
        if (eat("(")) {
            auto x = orx();
            if (!x) return x;
            if (!eat(")")) return fail_("')' expected");
            return x;
        }
        const char c = s[i];
        if (std::isdigit((unsigned char)c) || c == '.') {
            const char* b = s.c_str() + i;
            char* end = nullptr;
            double v = std::strtod(b, &end);
            if (end == b) return fail_("bad number");
            i += (size_t)(end - b);
            auto e = std::make_unique<Expr>();
            e->op = Expr::NUM;
            e->v = v;
            return e;
        }
        if (std::isalpha((unsigned char)c) || c == '_') {
            size_t b = i;
            while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_')) ++i;
            const std::string name = s.substr(b, i - b);
            auto e = std::make_unique<Expr>();
            for (size_t k = 0; k < vars.size(); ++k)
                if (name == vars[k]) {
                    e->op = Expr::VAR;
                    e->fn = (int)k;
                    return e;
                }
            if (name == "pi" || name == "e") {
                e->op = Expr::NUM;
                e->v = name == "pi" ? 3.14159265358979323846 : 2.71828182845904523536;
                return e;
            }
            static const char* f1[] = {"sin",  "cos",  "tan", "asin", "acos",  "atan", "sinh", "cosh",
                                       "tanh", "exp",  "log", "log10", "sqrt", "abs",  "floor", "ceil", "int"};
            static const char* f2[] = {"atan2", "pow", "min", "max"};
            int fn = -1, ar = 0;
            for (int k = 0; k < (int)(sizeof(f1) / sizeof(*f1)); ++k)
                if (name == f1[k]) fn = k, ar = 1;
            for (int k = 0; k < (int)(sizeof(f2) / sizeof(*f2)); ++k)
                if (name == f2[k]) fn = k, ar = 2;
            if (name == "if") fn = 0, ar = 3;
            if (fn < 0) return fail_("unknown name '" + name + "'");
            if (!eat("(")) return fail_("'(' expected after " + name);
            e->op = ar == 1 ? Expr::FN1 : (ar == 2 ? Expr::FN2 : Expr::IF);
            e->fn = fn;
            for (int k = 0; k < ar; ++k) {
                if (k && !eat(",")) return fail_("',' expected");
                auto x = orx();
                if (!x) return x;
                e->a.push_back(std::move(x));
            }
            if (!eat(")")) return fail_("')' expected");
            return e;
        }
        return fail_(std::string("unexpected '") + c + "'");
    }
};

double expr_eval(const Expr& e, const double* var)
{
    auto A = [&](int k) { return expr_eval(*e.a[k], var); };
    switch (e.op) {
    case Expr::NUM: return e.v;
    case Expr::VAR: return var[e.fn];
    case Expr::NEG: return -A(0);
    case Expr::NOT: return A(0) == 0 ? 1.0 : 0.0;
    case Expr::ADD: return A(0) + A(1);
    case Expr::SUB: return A(0) - A(1);
    case Expr::MUL: return A(0) * A(1);
    case Expr::DIV: return A(0) / A(1);
    case Expr::POW: return std::pow(A(0), A(1));
    case Expr::LT: return A(0) < A(1) ? 1.0 : 0.0;
    case Expr::LE: return A(0) <= A(1) ? 1.0 : 0.0;
    case Expr::GT: return A(0) > A(1) ? 1.0 : 0.0;
    case Expr::GE: return A(0) >= A(1) ? 1.0 : 0.0;
    case Expr::EQ: return A(0) == A(1) ? 1.0 : 0.0;
    case Expr::NE: return A(0) != A(1) ? 1.0 : 0.0;
    case Expr::AND: return (A(0) != 0 && A(1) != 0) ? 1.0 : 0.0;
    case Expr::OR: return (A(0) != 0 || A(1) != 0) ? 1.0 : 0.0;
    case Expr::IF: return A(0) != 0 ? A(1) : A(2);
    case Expr::FN1: {
        const double x = A(0);
        switch (e.fn) {
        case 0: return std::sin(x);
        case 1: return std::cos(x);
        case 2: return std::tan(x);
        case 3: return std::asin(x);
        case 4: return std::acos(x);
        case 5: return std::atan(x);
        case 6: return std::sinh(x);
        case 7: return std::cosh(x);
        case 8: return std::tanh(x);
        case 9: return std::exp(x);
        case 10: return std::log(x);
        case 11: return std::log10(x);
        case 12: return std::sqrt(x);
        case 13: return std::fabs(x);
        case 14: return std::floor(x);
        case 15: return std::ceil(x);
        default: return std::floor(x + 0.5);
        }
    }
    case Expr::FN2: {
        const double x = A(0), y = A(1);
        switch (e.fn) {
        case 0: return std::atan2(x, y);
        case 1: return std::pow(x, y);
        case 2: return x < y ? x : y;
        default: return x > y ? x : y;
        }
    }
    }
    return 0;
}

std::unique_ptr<Expr> parse_expr(const std::string& text, const std::vector<std::string>& vars, std::string& err)
{
    ExprParser P(text, vars);
    auto e = P.orx();
    P.ws();
    if (e && P.i != text.size()) e = P.fail_("unexpected text");
    err = P.err;
    // a failed operand leaves a node with a missing child behind: any
    // recorded error fails the whole parse
    if (!err.empty()) e.reset();
    return e;
}

// parse a custom signal f(t); on failure returns null and sets err
std::unique_ptr<Expr> parse_signal(const std::string& text, std::string& err)
{
    static const std::vector<std::string> vars = {"t"};
    return parse_expr(text, vars, err);
}

// parse a weight function of x, y, z, rho, r, a
std::unique_ptr<Expr> parse_weight(const std::string& text, std::string& err)
{
    static const std::vector<std::string> vars = {"x", "y", "z", "rho", "r", "a"};
    return parse_expr(text, vars, err);
}

double weight_eval(const ExcWeight& w, int n, const double* P)
{
    if (!w.f[n]) return 1.0;
    const double x = P[0], y = P[1], z = P[2];
    const double v[6] = {x, y, z, std::sqrt(x * x + y * y), std::sqrt(x * x + y * y + z * z), std::atan2(y, x)};
    return expr_eval(*w.f[n], v);
}
}  // namespace

static unsigned nyquist_num(double fmax, double dT)
{
    // CalcNyquistNum (tools/useful.cpp:30-36)
    if (fmax == 0) return 0xFFFFFFFFu;
    if (dT == 0) return 1;
    double T0 = 1 / fmax;
    return (unsigned)std::floor(T0 / 2 / dT);
}

int emsop::build_signal(unsigned nrts)
{
    sigV.clear();
    sigC.clear();
    switch (sigType) {
    case EMSOP_SIG_GAUSS: {
        sigLen = (unsigned)std::ceil(2.0 * 9.0 / (2.0 * kPI * fc) / dT);
        if (sigLen > nrts) sigLen = nrts;
        sigV.resize(sigLen);
        sigC.resize(sigLen);
        for (unsigned n = 0; n < sigLen; ++n) {
            double t = n * dT;
            sigV[n] = (float)(std::cos(2.0 * kPI * f0 * (t - 9.0 / (2.0 * kPI * fc))) * std::exp(-1 * std::pow(2.0 * kPI * fc * t / 3.0 - 3, 2)));
            t += 0.5 * dT;
            sigC[n] = (float)(std::cos(2.0 * kPI * f0 * (t - 9.0 / (2.0 * kPI * fc))) * std::exp(-1 * std::pow(2.0 * kPI * fc * t / 3.0 - 3, 2)));
        }
        nyquist = nyquist_num(f0 + fc, dT);
        break;
    }
    case EMSOP_SIG_SINUS: {
        sigLen = (unsigned)std::round(2.0 / f0 / dT);
        sigV.assign(sigLen, 0.0f);
        sigC.assign(sigLen, 0.0f);
        for (unsigned n = 1; n < sigLen; ++n) {
            double t = n * dT;
            sigV[n] = (float)std::sin(2.0 * kPI * f0 * t);
            t += 0.5 * dT;
            sigC[n] = (float)std::sin(2.0 * kPI * f0 * t);
        }
        nyquist = nyquist_num(f0, dT);
        break;
    }
    case EMSOP_SIG_DIRAC:
        sigLen = 2;
        sigV = {0.0f, 1.0f};
        sigC = {0.0f, 1.0f};
        nyquist = 1;
        break;
    case EMSOP_SIG_STEP:
        sigLen = 2;
        sigV = {1.0f, 1.0f};
        sigC = {1.0f, 1.0f};
        nyquist = 1;
        break;
    case EMSOP_SIG_CUSTOM: {
        // CalcCustomExcitation (excitation.cpp:219-252): nrts samples of f(t)
        std::string err;
        auto ex = parse_signal(sigFunc, err);
        if (!ex) return fail(EMSOP_E_ARG, "custom excitation: " + err);
        sigLen = nrts;
        sigV.resize(sigLen);
        sigC.resize(sigLen);
        for (unsigned n = 0; n < sigLen; ++n) {
            double t = n * dT;
            sigV[n] = (float)expr_eval(*ex, &t);
            t += 0.5 * dT;
            sigC[n] = (float)expr_eval(*ex, &t);
        }
        nyquist = nyquist_num(f0, dT);
        break;
    }
    default:
        return fail(EMSOP_E_STATE, "excitation signal not defined");
    }
    if (nyquist == 0) return fail(EMSOP_E_NUMERIC, "excitation setup failed (Nyquist number 0)");
    return 0;
}

// ---------------------------------------------------------------------------
// Operator::CalcECOperator (operator.cpp:986-1097)
// ---------------------------------------------------------------------------
int emsop::build(unsigned nrts)
{
    const bool verbose = std::getenv("EMSOP_VERBOSE") != nullptr;
    auto t0 = std::chrono::steady_clock::now();
    auto lap = [&](const char* what) {
        if (!verbose) return;
        auto t1 = std::chrono::steady_clock::now();
        std::fprintf(stderr, "emsop: %-18s %8.3f s\n", what, std::chrono::duration<double>(t1 - t0).count());
        t0 = t1;
    };
    for (int n = 0; n < 3; ++n)
        if (N[n] < 3) return fail(EMSOP_E_ARG, "at least 3 mesh lines per direction are required");
    if (sigType < 0) return fail(EMSOP_E_STATE, "excitation signal not defined");
    mesh_tables();
    try {
        int rc = calc_ec();
        if (rc) return rc;
        lap("calc_ec");
        double dt_auto = (tsMethod == 3) ? timestep_var3() : timestep_var1();
        lap("timestep");
        if (forcedDt > 0)
            dT = forcedDt;
        else
            dT = dt_auto;
        if (dT == 0 || dT >= 1e200) return fail(EMSOP_E_NUMERIC, "timestep is zero");
        dT *= tsFactor;
        period = (sigType == EMSOP_SIG_SINUS) ? 1 / f0 : 0;
        if (period > 0) {
            unsigned TS = (unsigned)std::ceil(period / dT);
            dT = period / TS;
        }
        const size_t tot = 3 * size();
        for (int k = 0; k < 4; ++k) coeff[k].alloc(tot);  // every entry is written below
        parallel_for(3 * N[0], nthreads, [&](unsigned a, unsigned b) {
            for (unsigned nx = a; nx < b; ++nx) {
                int n = nx / N[0];
                unsigned x = nx % N[0];
                for (unsigned y = 0; y < N[1]; ++y)
                    for (unsigned z = 0; z < N[2]; ++z) ec_to_coeff(n, idx(n, x, y, z));
            }
        });
        lap("coefficients");
        apply_bc();
        calc_pec();
        calc_pec_curves();
        lumped_rc();
        emsop_apply_magnetic_bc(this);
        lap("bc/pec/lumped");
        rc = build_signal(nrts);
        if (rc) return rc;
        // extensions in the reference's creation order (openems.cpp:1185-1243)
        build_excitation();
        lap("excitation");
        build_mur();
        lap("mur");
        build_upml();
        lap("upml");
        build_lorentz();
        build_rlc();
        lap("lorentz/rlc");
    } catch (const std::bad_alloc&) {
        return fail(EMSOP_E_NOMEM, "out of host memory while building the operator");
    }
    built = true;
    return 0;
}

// ---------------------------------------------------------------------------
// Lorentz / Drude and Debye materials (Operator_Ext_LorentzMaterial::
// BuildExtension, FDTD/extensions/operator_ext_lorentzmaterial.cpp:120-445).  Per order, the cells with an active component in x ->
// y -> z order; the material at an edge is the MATERIAL|METAL winner at its
// Yee point (primary edge centre for voltages, dual edge centre for
// currents).  Coefficients in double, stored as float.
// ---------------------------------------------------------------------------
void emsop::build_lorentz()
{
    lor.clear();
    unsigned order_max = 0;
    for (const auto& m : mats)
        if (m.disp >= 0) order_max = std::max(order_max, disps[m.disp].norder);
    if (!order_max) return;
    Ranked<MatBox> rk;
    rk.init(mats, [](const MatBox&) { return true; });
    const unsigned W = rk.W;
    // masks of the voltage (primary edge centre) and current (dual edge
    // centre) Yee points of each component
    std::vector<uint64_t> mv[3][3], mi[3][3];
    for (int n = 0; n < 3; ++n)
        for (int a = 0; a < 3; ++a) {
            mv[n][a] = rk.masks(a, a == n ? dualL[a] : lines[a]);
            mi[n][a] = rk.masks(a, a == n ? lines[a] : dualL[a]);
        }
    lor.resize(order_max);
    for (unsigned o = 0; o < order_max; ++o) {
        LorOrder& T = lor[o];
        std::vector<double> acc[6][3];
        unsigned p[3];
        for (p[0] = 0; p[0] < N[0]; ++p[0])
            for (p[1] = 0; p[1] < N[1]; ++p[1])
                for (p[2] = 0; p[2] < N[2]; ++p[2]) {
                    bool on = false;
                    double LD[3], RD[3], CL[3], CD[3], GD[3], LL[3];
                    for (int n = 0; n < 3; ++n) {
                        LD[n] = RD[n] = CL[n] = 0;
                        if (p[n] >= N[n] - 1) continue;                       // GetYeeCoords (operator.cpp:189-193)
                        const size_t q = idx(n, p[0], p[1], p[2]);
                        if (coeff[1][q] == 0) continue;
                        double P[3];
                        for (int a = 0; a < 3; ++a) P[a] = (a == n ? dualL[a] : lines[a])[p[a]];
                        const int r = rk.winner_at(&mv[n][0][(size_t)p[0] * W], &mv[n][1][(size_t)p[1] * W],
                                                   &mv[n][2][(size_t)p[2] * W], P);
                        if (r < 0 || rk.box[r]->disp < 0) continue;
                        const Dispersion& D = disps[rk.box[r]->disp];
                        if (D.debye) {
                            // CSPropDebyeMaterial (operator_ext_lorentzmaterial.cpp:259-275)
                            const double A = EL((n + 1) % 3, p[(n + 1) % 3], true) * EL((n + 2) % 3, p[(n + 2) % 3], true);
                            CL[n] = 8.85418781762e-12 * D.at(o, 0, n) * A / EL(n, p[n], false);
                            const double tr = D.at(o, 1, n);
                            if (CL[n] > 0 && tr > 0 && tr > 2.0 * dT) {
                                RD[n] = tr / CL[n];
                                on = true;
                                T.von = true;
                                T.vlon = true;
                            }
                            continue;
                        }
                        const double wp = D.at(o, 0, n) * 2 * kPI;
                        const double C = ec[0][q];
                        if (wp > 0 && C > 0) {
                            on = true;
                            T.von = true;
                            LD[n] = 1 / (wp * wp * C);
                        }
                        const double tr = D.at(o, 1, n);
                        if (tr > 0 && T.von) RD[n] = LD[n] / tr;
                        const double wl = D.at(o, 2, n) * 2 * kPI;
                        if (wl > 0 && LD[n] > 0) {
                            T.vlon = true;
                            CL[n] = 1 / (wl * wl * LD[n]);
                        }
                    }
                    for (int n = 0; n < 3; ++n) {
                        CD[n] = GD[n] = LL[n] = 0;
                        const int nP = (n + 1) % 3, nPP = (n + 2) % 3;
                        if (p[nP] >= N[nP] - 1 || p[nPP] >= N[nPP] - 1) continue;   // operator.cpp:194-200
                        const size_t q = idx(n, p[0], p[1], p[2]);
                        if (coeff[3][q] == 0) continue;
                        double P[3];
                        for (int a = 0; a < 3; ++a) P[a] = (a == n ? lines[a] : dualL[a])[p[a]];
                        const int r = rk.winner_at(&mi[n][0][(size_t)p[0] * W], &mi[n][1][(size_t)p[1] * W],
                                                   &mi[n][2][(size_t)p[2] * W], P);
                        if (r < 0 || rk.box[r]->disp < 0) continue;
                        const Dispersion& D = disps[rk.box[r]->disp];
                        if (D.debye) continue;     // no current-side Debye term
                        const double wp = D.at(o, 3, n) * 2 * kPI;
                        const double L = ec[2][q];
                        if (wp > 0 && L > 0) {
                            on = true;
                            T.ion = true;
                            CD[n] = 1 / (wp * wp * L);
                        }
                        const double tr = D.at(o, 4, n);
                        if (tr > 0 && T.ion) GD[n] = CD[n] / tr;
                        const double wl = D.at(o, 5, n) * 2 * kPI;
                        if (wl > 0 && CD[n] > 0) {
                            T.ilon = true;
                            LL[n] = 1 / (wl * wl * CD[n]);
                        }
                    }
                    if (!on) continue;
                    for (int n = 0; n < 3; ++n) {
                        T.pos[n].push_back(p[n]);
                        const size_t q = idx(n, p[0], p[1], p[2]);
                        const double VI = coeff[1][q], IV = coeff[3][q];
                        if (LD[n] > 0) {
                            acc[0][n].push_back((2.0 * LD[n] - dT * RD[n]) / (2.0 * LD[n] + dT * RD[n]));
                            acc[1][n].push_back(dT / (LD[n] + dT * RD[n] / 2.0) * VI);
                        } else if (RD[n] > 0 && CL[n] > 0) {
                            // Debye: R_D C_L = tau
                            acc[0][n].push_back((2.0 * dT - RD[n] * CL[n]) / (CL[n] * RD[n]));
                            acc[1][n].push_back(2.0 / RD[n] * VI);
                        } else {
                            acc[0][n].push_back(1);
                            acc[1][n].push_back(0);
                        }
                        if (CD[n] > 0) {
                            acc[2][n].push_back((2.0 * CD[n] - dT * GD[n]) / (2.0 * CD[n] + dT * GD[n]));
                            acc[3][n].push_back(dT / (CD[n] + dT * GD[n] / 2.0) * IV);
                        } else {
                            acc[2][n].push_back(1);
                            acc[3][n].push_back(0);
                        }
                        acc[4][n].push_back(CL[n] > 0 ? dT / CL[n] / VI : 0.0);
                        acc[5][n].push_back(LL[n] > 0 ? dT / LL[n] / IV : 0.0);
                    }
                }
        for (int k = 0; k < 6; ++k)
            for (int n = 0; n < 3; ++n) {
                T.t[k][n].resize(acc[k][n].size());
                for (size_t i = 0; i < acc[k][n].size(); ++i) T.t[k][n][i] = (float)acc[k][n][i];
            }
    }
}

// ---------------------------------------------------------------------------
// Lumped RLC elements (Operator_Ext_LumpedRLC::BuildExtension,
// FDTD/extensions/operator_ext_lumpedRLC.cpp:112-548): parallel elements
// with L > 0 and series elements, in scene order.  Rewrites EC_C / EC_G and
// the coefficients of their edges and builds the ADE tables.
// ---------------------------------------------------------------------------
void emsop::build_rlc()
{
    constexpr double kZFact = 20.0;     // LUMPED_RLC_Z_FACT (operator_ext_lumpedRLC.h:29)
    rlc = RlcTab{};
    const double fMax = f0 + fc;
    for (const auto& le : lumped) {
        double C = le.C, R = le.R, L = le.L;
        if (C < 0.0) C = 0.0;
        if (R < 0.0) R = 0;
        if (L < 0.0) L = 0;
        if (std::isnan(C)) C = 0.0;
        if (std::isnan(L)) L = 0.0;
        if (std::isnan(R)) R = 0.0;
        const bool parallel = le.parallel;
        // IsLElumpedRLC (operator_ext_lumpedRLC.cpp:550-561) reads the raw inductance
        if (!((parallel && le.L > 0.0) || !parallel)) continue;
        const int dir = le.dir;
        if (dir < 0 || dir > 2) continue;
        const int d1 = (dir + 1) % 3, d2 = (dir + 2) % 3;
        unsigned s[3], e[3];
        const int dim = snap_box(le.lo, le.hi, false, s, e, nullptr, nullptr);
        if (dim <= 0) continue;
        if (s[dir] == e[dir]) continue;
        const unsigned n0 = e[dir] - s[dir], n1 = e[d1] - s[d1] + 1, n2 = e[d2] - s[d2] + 1;
        const unsigned Npar = n1 * n2;
        const double dL = L * Npar / n0, dR = R * Npar / n0;
        double dG = (1 / R) * n0 / Npar;
        double dC = C * n0 / Npar;
        double ib0 = 2.0 * dT * dC / (4.0 * dL * dC + 2.0 * dT * dR * dC + dT * dT);
        double b1 = (dT * dT - 4.0 * dL * dC) / (dT * dC);
        double b2 = (4.0 * dL * dC - 2.0 * dT * dR * dC + dT * dT) / (2.0 * dT * dC);
        if (!parallel && C == 0) {
            ib0 = dT / (2.0 * dL + dT * dR);
            b1 = -4.0 * dL / dT;
            b2 = (2.0 * dL - dT * dR) / dT;
        }
        if (parallel && R == 0.0) dG = 0.0;
        unsigned p[3];
        for (p[dir] = s[dir]; p[dir] < e[dir]; ++p[dir])
            for (p[d1] = s[d1]; p[d1] <= e[d1]; ++p[d1])
                for (p[d2] = s[d2]; p[d2] <= e[d2]; ++p[d2]) {
                    const size_t q = idx(dir, p[0], p[1], p[2]);
                    double v[9];
                    if (parallel) {
                        // a set capacitor replaces the edge's own; otherwise the
                        // edge's capacitance is taken (and kept for the later edges)
                        if (dC > 0)
                            ec[0][q] = (float)dC;
                        else
                            dC = ec[0][q];
                        v[0] = (dT / dC) / (1.0 + dT * dG / (2.0 * dC));
                        if (R >= 0) ec[1][q] = (float)dG;
                        v[1] = L > 0 ? dT / dL : 0.0;
                        if (dC == 0) {
                            double Cd = ec[0][q];
                            const double Zmin = std::max(dR, 2 * kPI * fMax * dL);
                            const double Zcd_min = 1.0 / (2.0 * kPI * fMax * Cd);
                            if (Zcd_min < kZFact * Zmin) {
                                Cd = 1.0 / (2 * kPI * fMax * Zmin * kZFact);
                                ec[0][q] = (float)Cd;
                            }
                        }
                        v[2] = 1.0;
                        v[3] = v[4] = v[5] = v[6] = v[7] = v[8] = 0.0;
                    } else {
                        ec[1][q] = 0.0f;
                        float Cd = ec[0][q];
                        // pow(v, 2) of the reference, which GCC folds to v * v
                        double Zmin;
                        if (dC) {
                            const double X = 2 * kPI * fMax * dL - 1.0 / (dC * 2 * kPI * fMax);
                            Zmin = std::sqrt(dR * dR + X * X);
                        } else {
                            const double X = 2 * kPI * fMax * dL;
                            Zmin = std::sqrt(dR * dR + X * X);
                        }
                        const double Zcd_min = 1.0 / (2.0 * kPI * fMax * Cd);
                        if (Zcd_min < kZFact * Zmin) {
                            Cd = (float)(1.0 / (2 * kPI * fMax * Zmin * kZFact));
                            ec[0][q] = Cd;
                        }
                        v[0] = 0.0;
                        v[1] = 0.0;
                        v[3] = 0.5 * dT * ib0 / Cd;
                        v[4] = 0.5 * dT * (b1 * ib0 - 1.0) / Cd;
                        v[5] = 0.5 * dT * b2 * ib0 / Cd;
                        v[2] = 1.0 / (1.0 + 0.5 * dT * ib0 / Cd);
                        v[6] = ib0;
                        v[7] = b1;
                        v[8] = b2;
                    }
                    ec_to_coeff(dir, q);
                    rlc.dir.push_back(dir);
                    for (int k = 0; k < 9; ++k) rlc.c[k].push_back((float)v[k]);
                    for (int a = 0; a < 3; ++a) rlc.pos[a].push_back(p[a]);
                }
        if (le.caps) {
            for (p[d1] = s[d1]; p[d1] <= e[d1]; ++p[d1])
                for (p[d2] = s[d2]; p[d2] <= e[d2]; ++p[d2])
                    for (unsigned pl : {s[dir], e[dir]}) {
                        p[dir] = pl;
                        if (p[d1] < e[d1]) {
                            coeff[0][idx(d1, p[0], p[1], p[2])] = 0;
                            coeff[1][idx(d1, p[0], p[1], p[2])] = 0;
                        }
                        if (p[d2] < e[d2]) {
                            coeff[0][idx(d2, p[0], p[1], p[2])] = 0;
                            coeff[1][idx(d2, p[0], p[1], p[2])] = 0;
                        }
                    }
        }
    }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* emsop_last_error(void) { return g_err.c_str(); }

int emsop_create(emsop** out)
{
    if (!out) return fail(EMSOP_E_ARG, "null output pointer");
    try {
        *out = new emsop();
    } catch (const std::bad_alloc&) {
        return fail(EMSOP_E_NOMEM, "out of memory");
    }
    return 0;
}

void emsop_destroy(emsop* h) { delete h; }

int emsop_set_grid(emsop* h, const double* x, unsigned nx, const double* y, unsigned ny, const double* z, unsigned nz,
                   double delta_unit)
{
    if (!h || !x || !y || !z) return fail(EMSOP_E_ARG, "null argument");
    if (nx < 3 || ny < 3 || nz < 3) return fail(EMSOP_E_ARG, "at least 3 mesh lines per direction are required");
    if (delta_unit <= 0) return fail(EMSOP_E_ARG, "grid delta unit must be > 0");
    const double* src[3] = {x, y, z};
    unsigned cnt[3] = {nx, ny, nz};
    for (int n = 0; n < 3; ++n) {
        h->lines[n].assign(src[n], src[n] + cnt[n]);
        for (unsigned i = 1; i < cnt[n]; ++i)
            if (!(h->lines[n][i] > h->lines[n][i - 1])) return fail(EMSOP_E_ARG, "mesh lines must be strictly increasing");
        h->N[n] = cnt[n];
    }
    h->gd = delta_unit;
    h->built = false;
    return 0;
}

int emsop_set_background(emsop* h, double eps_r, double mue_r, double kappa, double sigma)
{
    if (!h) return fail(EMSOP_E_ARG, "null handle");
    h->bg[0] = eps_r;
    h->bg[1] = mue_r;
    h->bg[2] = kappa;
    h->bg[3] = sigma;
    return 0;
}

static int next_order(emsop* h) { return (int)(h->mats.size() + h->lumped.size() + h->excs.size()); }

int emsop_add_material(emsop* h, const double eps[3], const double mue[3], const double kappa[3], const double sigma[3],
                       int priority, const double start[3], const double stop[3])
{
    if (!h || !eps || !mue || !kappa || !sigma || !start || !stop) return fail(EMSOP_E_ARG, "null argument");
    MatBox b;
    for (int n = 0; n < 3; ++n) {
        b.lo[n] = start[n];
        b.hi[n] = stop[n];
        b.eps[n] = eps[n];
        b.mue[n] = mue[n];
        b.kappa[n] = kappa[n];
        b.sigma[n] = sigma[n];
        if (!(mue[n] > 0) || !(eps[n] > 0)) return fail(EMSOP_E_ARG, "eps_r and mue_r must be > 0");
    }
    b.prio = priority;
    b.order = next_order(h);
    b.metal = false;
    h->mats.push_back(b);
    return 0;
}

int emsop_set_dispersion(emsop* h, unsigned norder, const double* params)
{
    if (!h || (norder && !params)) return fail(EMSOP_E_ARG, "null argument");
    if (h->mats.empty() || h->mats.back().metal || h->mats.back().grid >= 0)
        return fail(EMSOP_E_STATE, "dispersion needs a material box added just before");
    if (norder == 0 || norder > 64) return fail(EMSOP_E_ARG, "dispersion order must be 1..64");
    Dispersion d;
    d.norder = norder;
    d.p.assign(params, params + (size_t)norder * 18);
    for (double v : d.p)
        if (!(v >= 0)) return fail(EMSOP_E_ARG, "dispersion frequencies and times must be >= 0");
    h->mats.back().disp = (int)h->disps.size();
    h->disps.push_back(std::move(d));
    return 0;
}

int emsop_set_debye(emsop* h, unsigned norder, const double* params)
{
    if (!h || !params) return fail(EMSOP_E_ARG, "null argument");
    if (h->mats.empty() || h->mats.back().metal || h->mats.back().grid >= 0)
        return fail(EMSOP_E_STATE, "dispersion needs a material box added just before");
    if (norder == 0 || norder > 64) return fail(EMSOP_E_ARG, "dispersion order must be 1..64");
    Dispersion d;
    d.norder = norder;
    d.debye = true;
    d.p.assign((size_t)norder * 18, 0.0);
    for (unsigned o = 0; o < norder; ++o)
        for (int k = 0; k < 2; ++k)
            for (int n = 0; n < 3; ++n) {
                const double v = params[((size_t)o * 2 + k) * 3 + n];
                if (!(v >= 0)) return fail(EMSOP_E_ARG, "Debye eps delta and relaxation time must be >= 0");
                d.p[((size_t)o * 6 + k) * 3 + n] = v;
            }
    h->mats.back().disp = (int)h->disps.size();
    h->disps.push_back(std::move(d));
    return 0;
}

int emsop_add_material_grid(emsop* h, const unsigned cell0[3], const unsigned ncells[3], const double* eps_r,
                            const double* kappa, double mue_r, double sigma, int priority)
{
    if (!h || !cell0 || !ncells || !eps_r) return fail(EMSOP_E_ARG, "null argument");
    for (int n = 0; n < 3; ++n) {
        if (h->lines[n].size() < 2) return fail(EMSOP_E_STATE, "set the grid before adding a material grid");
        if (ncells[n] == 0 || (size_t)cell0[n] + ncells[n] >= h->lines[n].size())
            return fail(EMSOP_E_ARG, "material grid outside the mesh");
    }
    if (!(mue_r > 0)) return fail(EMSOP_E_ARG, "mue_r must be > 0");
    MatGrid G;
    size_t cnt = 1;
    for (int n = 0; n < 3; ++n) {
        G.i0[n] = cell0[n];
        G.n[n] = ncells[n];
        cnt *= ncells[n];
    }
    G.eps.assign(eps_r, eps_r + cnt);
    G.kappa.assign(cnt, 0.0);
    if (kappa) G.kappa.assign(kappa, kappa + cnt);
    for (size_t q = 0; q < cnt; ++q)
        if (!(G.eps[q] > 0)) return fail(EMSOP_E_ARG, "eps_r must be > 0");
    MatBox b;
    for (int n = 0; n < 3; ++n) {
        b.lo[n] = h->lines[n][cell0[n]];
        b.hi[n] = h->lines[n][cell0[n] + ncells[n]];
        b.eps[n] = 1.0;
        b.mue[n] = mue_r;
        b.kappa[n] = 0.0;
        b.sigma[n] = sigma;
    }
    b.prio = priority;
    b.order = next_order(h);
    b.metal = false;
    b.grid = (int)h->grids.size();
    h->grids.push_back(std::move(G));
    h->mats.push_back(b);
    return 0;
}

int emsop_eps_weights(const emsop* h, int n, const unsigned pos[3], unsigned cells[12], double dC[4])
{
    if (!h || !pos || !cells || !dC) return fail(EMSOP_E_ARG, "null argument");
    if (n < 0 || n > 2) return fail(EMSOP_E_ARG, "bad component");
    if (h->cellConst) return fail(EMSOP_E_STATE, "weights are defined for quarter-cell averaging");
    for (int a = 0; a < 3; ++a)
        if (pos[a] >= h->N[a]) return fail(EMSOP_E_ARG, "edge outside the mesh");
    const int nP = (n + 1) % 3, nPP = (n + 2) % 3;
    const int iP = pos[nP], iPP = pos[nPP];
    const int sP[4] = {0, -1, 0, -1}, sPP[4] = {0, 0, -1, -1};
    double A[4], area = 0;
    for (int q = 0; q < 4; ++q) {
        A[q] = h->Wp(nP, iP + sP[q]) * h->Wp(nPP, iPP + sPP[q]);
        area += A[q];
    }
    const double delta = h->EL(n, pos[n], false);
    const double earea = h->EL(nP, pos[nP], true) * h->EL(nPP, pos[nPP], true);
    for (int q = 0; q < 4; ++q) {
        int c[3];
        c[n] = (int)pos[n];
        c[nP] = iP + sP[q];
        c[nPP] = iPP + sPP[q];
        for (int a = 0; a < 3; ++a) cells[3 * q + a] = c[a] < 0 ? 0xFFFFFFFFu : (unsigned)c[a];
        // C = eps0 * sum_q eps_q A_q / area * earea / delta (AverageMatQuarterCell + Calc_ECPos)
        dC[q] = (delta && area > 0) ? kEPS0 * A[q] / area * earea / delta : 0.0;
    }
    return 0;
}

int emsop_add_metal(emsop* h, int priority, const double start[3], const double stop[3])
{
    if (!h || !start || !stop) return fail(EMSOP_E_ARG, "null argument");
    MatBox b;
    for (int n = 0; n < 3; ++n) {
        b.lo[n] = start[n];
        b.hi[n] = stop[n];
        b.eps[n] = b.mue[n] = 1;
        b.kappa[n] = b.sigma[n] = 0;
    }
    b.prio = priority;
    b.order = next_order(h);
    b.metal = true;
    h->mats.push_back(b);
    return 0;
}

int emsop_add_lumped(emsop* h, int dir, double R, double C, double L, int caps, int parallel, int priority,
                     const double start[3], const double stop[3])
{
    if (!h || !start || !stop) return fail(EMSOP_E_ARG, "null argument");
    if (dir < 0 || dir > 2) return fail(EMSOP_E_ARG, "lumped element direction must be 0, 1 or 2");
    LumpedBox b;
    for (int n = 0; n < 3; ++n) {
        b.lo[n] = start[n];
        b.hi[n] = stop[n];
    }
    b.dir = dir;
    b.R = R;
    b.C = C;
    b.L = L;
    b.caps = caps != 0;
    b.parallel = parallel != 0;
    b.prio = priority;
    b.order = next_order(h);
    h->lumped.push_back(b);
    return 0;
}

int emsop_add_excitation(emsop* h, int type, const double val[3], double delay, int priority, const double start[3],
                         const double stop[3])
{
    if (!h || !val || !start || !stop) return fail(EMSOP_E_ARG, "null argument");
    if (type < 0 || type > 3) return fail(EMSOP_E_ARG, "excitation type must be 0..3");
    ExcBox b;
    for (int n = 0; n < 3; ++n) {
        b.lo[n] = start[n];
        b.hi[n] = stop[n];
        b.val[n] = val[n];
    }
    b.type = type;
    b.delay = delay;
    b.prio = priority;
    b.order = next_order(h);
    h->excs.push_back(b);
    h->last_exc = 1;
    return 0;
}

static int check_cylinder(const double start[3], const double stop[3], double radius, double shell_width)
{
    if (!(radius > 0) || !(shell_width >= 0)) return fail(EMSOP_E_ARG, "cylinder radius must be > 0 and shell width >= 0");
    if (start[0] == stop[0] && start[1] == stop[1] && start[2] == stop[2])
        return fail(EMSOP_E_ARG, "cylinder axis has zero length");
    return 0;
}

int emsop_add_metal_cylinder(emsop* h, int priority, const double start[3], const double stop[3], double radius,
                             double shell_width)
{
    if (!h || !start || !stop) return fail(EMSOP_E_ARG, "null argument");
    if (int rc = check_cylinder(start, stop, radius, shell_width)) return rc;
    MatBox b;
    b.set_cylinder(start, stop, radius, shell_width);
    for (int n = 0; n < 3; ++n) {
        b.eps[n] = b.mue[n] = 1;
        b.kappa[n] = b.sigma[n] = 0;
    }
    b.prio = priority;
    b.order = next_order(h);
    b.metal = true;
    h->mats.push_back(b);
    return 0;
}

int emsop_add_excitation_cylinder(emsop* h, int type, const double val[3], double delay, int priority,
                                  const double start[3], const double stop[3], double radius, double shell_width)
{
    if (!h || !val || !start || !stop) return fail(EMSOP_E_ARG, "null argument");
    if (type < 0 || type > 3) return fail(EMSOP_E_ARG, "excitation type must be 0..3");
    if (int rc = check_cylinder(start, stop, radius, shell_width)) return rc;
    ExcBox b;
    b.set_cylinder(start, stop, radius, shell_width);
    for (int n = 0; n < 3; ++n) b.val[n] = val[n];
    b.type = type;
    b.delay = delay;
    b.prio = priority;
    b.order = next_order(h);
    h->excs.push_back(b);
    h->last_exc = 1;
    return 0;
}

int emsop_eval_function(const char* text, const char* vars, unsigned nvar, const double* values, unsigned npts,
                        double* out)
{
    if (!text || !vars || (npts && (!values || !out))) return fail(EMSOP_E_ARG, "null argument");
    std::vector<std::string> names;
    std::string cur;
    for (const char* c = vars;; ++c) {
        if (*c == ',' || *c == 0) {
            names.push_back(cur);
            cur.clear();
            if (!*c) break;
        } else if (!std::isspace((unsigned char)*c)) {
            cur += *c;
        }
    }
    if (names.size() != nvar) return fail(EMSOP_E_ARG, "variable list does not hold nvar names");
    std::string err;
    auto e = parse_expr(text, names, err);
    if (!e) return fail(EMSOP_E_ARG, "function: " + err);
    for (unsigned k = 0; k < npts; ++k) out[k] = expr_eval(*e, values + (size_t)k * nvar);
    return 0;
}

int emsop_set_excitation_weight(emsop* h, const char* const weight[3])
{
    if (!h || !weight) return fail(EMSOP_E_ARG, "null argument");
    if (h->last_exc == 0) return fail(EMSOP_E_STATE, "no excitation primitive to weight");
    auto w = std::make_shared<ExcWeight>();
    for (int n = 0; n < 3; ++n) {
        if (!weight[n]) continue;   // no function: weight 1
        std::string err;
        auto e = parse_weight(weight[n], err);
        if (!e) return fail(EMSOP_E_ARG, std::string("excitation weight ") + "xyz"[n] + ": " + err);
        w->f[n] = std::shared_ptr<const Expr>(std::move(e));
    }
    if (h->last_exc == 1) h->excs.back().weight = w;
    else h->curves.back().weight = w;
    return 0;
}

int emsop_add_metal_curve(emsop* h, unsigned npts, const double* pts)
{
    if (!h || !pts) return fail(EMSOP_E_ARG, "null argument");
    if (npts < 2) return fail(EMSOP_E_ARG, "a curve needs at least two points");
    Curve c;
    c.pts.assign(pts, pts + 3 * (size_t)npts);
    h->curves.push_back(c);
    return 0;
}

int emsop_add_excitation_curve(emsop* h, int type, const double val[3], double delay, unsigned npts, const double* pts)
{
    if (!h || !val || !pts) return fail(EMSOP_E_ARG, "null argument");
    if (type < 0 || type > 3) return fail(EMSOP_E_ARG, "excitation type must be 0..3");
    if (npts < 2) return fail(EMSOP_E_ARG, "a curve needs at least two points");
    Curve c;
    c.pts.assign(pts, pts + 3 * (size_t)npts);
    c.exc = true;
    c.type = type;
    for (int n = 0; n < 3; ++n) c.val[n] = val[n];
    c.delay = delay;
    h->curves.push_back(c);
    h->last_exc = 2;
    return 0;
}

int emsop_set_boundary(emsop* h, const int bc[6], const unsigned pml_size[6], const double mur_vph[6])
{
    if (!h || !bc) return fail(EMSOP_E_ARG, "null argument");
    for (int k = 0; k < 6; ++k) {
        if (bc[k] < -1 || bc[k] > 3) return fail(EMSOP_E_ARG, "boundary type must be -1..3");
        h->bc[k] = bc[k];
        if (pml_size) h->pml[k] = pml_size[k];
        if (mur_vph) h->vph[k] = mur_vph[k];
    }
    return 0;
}

int emsop_set_signal(emsop* h, int type, double f0, double fc)
{
    if (!h) return fail(EMSOP_E_ARG, "null handle");
    if (type < 0 || type > 3) return fail(EMSOP_E_ARG, "unknown signal type");
    if (type == EMSOP_SIG_GAUSS && !(fc > 0)) return fail(EMSOP_E_ARG, "gaussian pulse needs fc > 0");
    if (type == EMSOP_SIG_SINUS && !(f0 > 0)) return fail(EMSOP_E_ARG, "sinusoid needs f0 > 0");
    h->sigType = type;
    h->f0 = f0;
    h->fc = fc;
    return 0;
}

int emsop_set_custom_signal(emsop* h, const char* function, double f0)
{
    if (!h || !function) return fail(EMSOP_E_ARG, "null argument");
    std::string err;
    if (!parse_signal(function, err)) return fail(EMSOP_E_ARG, "custom excitation: " + err);
    h->sigType = EMSOP_SIG_CUSTOM;
    h->sigFunc = function;
    h->f0 = f0;
    h->fc = 0;
    return 0;
}

int emsop_set_timestep(emsop* h, int method, double forced_dt, double factor)
{
    if (!h) return fail(EMSOP_E_ARG, "null handle");
    if (method != 1 && method != 3) return fail(EMSOP_E_ARG, "timestep method must be 1 or 3");
    h->tsMethod = method;
    h->forcedDt = forced_dt;
    if (factor > 0 && factor <= 1) h->tsFactor = factor;  // Operator::SetTimestepFactor ignores others
    return 0;
}

int emsop_set_cell_constant_material(emsop* h, int enable)
{
    if (!h) return fail(EMSOP_E_ARG, "null handle");
    h->cellConst = enable != 0;
    return 0;
}

int emsop_set_threads(emsop* h, int nthreads)
{
    if (!h) return fail(EMSOP_E_ARG, "null handle");
    h->nthreads = nthreads;
    return 0;
}

int emsop_build(emsop* h, unsigned nrts)
{
    if (!h) return fail(EMSOP_E_ARG, "null handle");
    if (nrts == 0) return fail(EMSOP_E_ARG, "number of timesteps must be > 0");
    return h->build(nrts);
}

int emsop_num_lines(const emsop* h, unsigned out[3])
{
    if (!h || !out) return fail(EMSOP_E_ARG, "null argument");
    for (int n = 0; n < 3; ++n) out[n] = h->N[n];
    return 0;
}

double emsop_timestep(const emsop* h) { return h ? h->dT : 0; }
unsigned emsop_nyquist(const emsop* h) { return h ? h->nyquist : 0; }
double emsop_signal_period(const emsop* h) { return h ? h->period : 0; }
unsigned emsop_signal_length(const emsop* h) { return h ? h->sigLen : 0; }
int emsop_signal_period_steps(const emsop* h) { return (h && h->period > 0) ? (int)(h->period / h->dT) : 0; }
const float* emsop_signal_volt(const emsop* h) { return (h && h->built) ? h->sigV.data() : nullptr; }
const float* emsop_signal_curr(const emsop* h) { return (h && h->built) ? h->sigC.data() : nullptr; }

const float* emsop_coeff(const emsop* h, int which)
{
    if (!h || !h->built || which < 0 || which > 3 || h->coeff[which].empty()) return nullptr;
    return h->coeff[which].data();
}

const float* emsop_ec(const emsop* h, int which)
{
    if (!h || !h->built || which < 0 || which > 3 || h->ec[which].empty()) return nullptr;
    return h->ec[which].data();
}

int emsop_effmat_at(const emsop* h, int ny, const unsigned pos[3], double out[4])
{
    if (!h || !h->built || ny < 0 || ny > 2 || !pos || !out) return fail(EMSOP_E_ARG, "bad argument");
    for (int n = 0; n < 3; ++n)
        if (pos[n] >= h->N[n]) return fail(EMSOP_E_ARG, "position outside the mesh");
    h->effmat_at(ny, pos, out);
    return 0;
}

unsigned emsop_exc_count(const emsop* h, int curr)
{
    if (!h || !h->built) return 0;
    return (unsigned)h->exc[curr ? 1 : 0].dir.size();
}

int emsop_exc_get(const emsop* h, int curr, unsigned* idx, unsigned short* dir, float* amp, unsigned* delay)
{
    if (!h || !h->built) return fail(EMSOP_E_STATE, "operator not built");
    const ExcList& e = h->exc[curr ? 1 : 0];
    const size_t K = e.dir.size();
    for (size_t k = 0; k < K; ++k) {
        if (idx)
            for (int a = 0; a < 3; ++a) idx[a * K + k] = e.idx[a][k];
        if (dir) dir[k] = e.dir[k];
        if (amp) amp[k] = e.amp[k];
        if (delay) delay[k] = e.delay[k];
    }
    return 0;
}

unsigned emsop_upml_count(const emsop* h) { return (h && h->built) ? (unsigned)h->upml.size() : 0; }

int emsop_upml_get(const emsop* h, unsigned k, unsigned start[3], unsigned num[3], const float* arrays[6])
{
    if (!h || !h->built || k >= h->upml.size()) return fail(EMSOP_E_ARG, "bad UPML slab index");
    const Slab& s = h->upml[k];
    for (int n = 0; n < 3; ++n) {
        if (start) start[n] = s.start[n];
        if (num) num[n] = s.num[n];
    }
    if (arrays)
        for (int a = 0; a < 6; ++a) arrays[a] = s.a[a].data();
    return 0;
}

unsigned emsop_lorentz_orders(const emsop* h) { return h && h->built ? (unsigned)h->lor.size() : 0; }

int emsop_lorentz_get(const emsop* h, unsigned order, unsigned* count, int on[4], const unsigned* pos[3],
                      const float* tabs[18])
{
    if (!h || !count || !on || !pos || !tabs) return fail(EMSOP_E_ARG, "null argument");
    if (!h->built) return fail(EMSOP_E_STATE, "operator not built");
    if (order >= h->lor.size()) return fail(EMSOP_E_ARG, "dispersion order out of range");
    const LorOrder& T = h->lor[order];
    *count = (unsigned)T.pos[0].size();
    on[0] = T.von;
    on[1] = T.ion;
    on[2] = T.vlon;
    on[3] = T.ilon;
    for (int n = 0; n < 3; ++n) pos[n] = T.pos[n].data();
    for (int k = 0; k < 6; ++k)
        for (int n = 0; n < 3; ++n) tabs[3 * k + n] = T.t[k][n].data();
    return 0;
}

unsigned emsop_rlc_count(const emsop* h) { return h && h->built ? (unsigned)h->rlc.dir.size() : 0; }

int emsop_rlc_get(const emsop* h, const unsigned* pos[3], const int** dir, const float* coef[9])
{
    if (!h || !pos || !dir || !coef) return fail(EMSOP_E_ARG, "null argument");
    if (!h->built) return fail(EMSOP_E_STATE, "operator not built");
    for (int n = 0; n < 3; ++n) pos[n] = h->rlc.pos[n].data();
    *dir = h->rlc.dir.data();
    for (int k = 0; k < 9; ++k) coef[k] = h->rlc.c[k].data();
    return 0;
}

unsigned emsop_mur_count(const emsop* h) { return (h && h->built) ? (unsigned)h->mur.size() : 0; }

int emsop_mur_get(const emsop* h, unsigned k, int* ny, unsigned* line, unsigned* line_shift, unsigned* start_ts,
                  const float** cP, const float** cPP)
{
    if (!h || !h->built || k >= h->mur.size()) return fail(EMSOP_E_ARG, "bad Mur face index");
    const MurFace& f = h->mur[k];
    if (ny) *ny = f.ny;
    if (line) *line = f.line;
    if (line_shift) *line_shift = f.shift;
    if (start_ts) *start_ts = f.start_ts;
    if (cP) *cP = f.cP.data();
    if (cPP) *cPP = f.cPP.data();
    return 0;
}

double emsop_disc_line(const emsop* h, int n, unsigned pos, int dual)
{
    if (!h || n < 0 || n > 2 || pos >= h->N[n] || h->dualL[n].empty()) return 0;
    return dual ? h->dualL[n][pos] : h->lines[n][pos];
}

double emsop_edge_length(const emsop* h, int n, unsigned pos, int dual)
{
    if (!h || n < 0 || n > 2 || pos >= h->N[n] || h->primD[n].empty()) return 0;
    return h->EL(n, pos, dual != 0);
}

int emsop_snap_to_mesh(const emsop* h, const double coord[3], int dual, unsigned out[3], int inside[3])
{
    if (!h || !coord || !out || h->dualL[0].empty()) return fail(EMSOP_E_ARG, "bad argument or mesh not prepared");
    int ok = 1;
    for (int n = 0; n < 3; ++n) {
        bool in;
        out[n] = h->snap_line(n, coord[n], in, dual != 0);
        if (inside) inside[n] = in;
        ok &= in;
    }
    return ok;
}

int emsop_snap_box(const emsop* h, const double start[3], const double stop[3], int dual, unsigned us[3], unsigned ue[3],
                   int start_in[3], int stop_in[3])
{
    if (!h || !start || !stop || !us || !ue || h->dualL[0].empty()) return fail(EMSOP_E_ARG, "bad argument or mesh not prepared");
    bool si[3] = {false, false, false}, ei[3] = {false, false, false};
    int dim = h->snap_box(start, stop, dual != 0, us, ue, si, ei);
    for (int n = 0; n < 3; ++n) {
        if (start_in) start_in[n] = si[n];
        if (stop_in) stop_in[n] = ei[n];
    }
    return dim;
}

int emsop_find_path(const emsop* h, const double start[3], const double stop[3], unsigned max_edges, unsigned* pos,
                    int* dir, unsigned* count)
{
    if (!h || !start || !stop || !count || h->dualL[0].empty()) return fail(EMSOP_E_ARG, "bad argument or mesh not prepared");
    std::vector<PathEdge> path;
    h->find_path(start, stop, path);
    *count = (unsigned)path.size();
    if (path.size() > max_edges) return fail(EMSOP_E_ARG, "path longer than max_edges");
    for (size_t t = 0; t < path.size(); ++t) {
        if (pos)
            for (int a = 0; a < 3; ++a) pos[3 * t + a] = path[t].pos[a];
        if (dir) dir[t] = path[t].dir;
    }
    return 0;
}

void emsop_release_build_arrays(emsop* h)
{
    if (!h) return;
    for (int k = 0; k < 4; ++k) h->ec[k].clear();
}

}  // extern "C"
