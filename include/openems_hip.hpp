/*
 * openems_hip.hpp — C++ host classes for the MI355X engine, with the names and
 * argument meaning of openEMS's Operator and Engine on the timestep path.
 *
 * Header-only, over the two C ABIs (ems_operator.h: the operator build of box
 * scenes; ems_hip.h: the device engine).  The reference classes they mirror:
 *
 *   Operator_HIP::CalcECOperator        Operator::CalcECOperator          FDTD/operator.cpp:986-1097
 *   Operator_HIP::GetTimestep           Operator_Base::GetTimestep        Common/operator_base.h:42
 *   Operator_HIP::GetNumberOfLines      Operator_Base::GetNumberOfLines   Common/operator_base.h:51
 *   Operator_HIP::GetDiscLine/GetEdgeLength
 *                                       Operator::GetDiscLine / GetEdgeLength
 *                                                                         FDTD/operator.cpp:144-215
 *   Operator_HIP::GetVV/GetVI/GetII/GetIV
 *                                       Operator::GetVV / GetVI / GetII / GetIV
 *                                                                         FDTD/operator.h:64-86
 *   Operator_HIP::CreateEngine          Operator::CreateEngine            FDTD/operator.h:56, FDTD/operator.cpp:68-72
 *   Engine_HIP::New / Init / Reset      Engine::New / Init / Reset        FDTD/engine.cpp:27-108
 *   Engine_HIP::IterateTS               Engine::IterateTS                 FDTD/engine.cpp:267-286
 *   Engine_HIP::GetNumberOfTimesteps    Engine::GetNumberOfTimesteps      FDTD/engine.h:50
 *   Engine_HIP::GetVolt/GetCurr/SetVolt/SetCurr
 *                                       Engine::GetVolt / GetCurr / SetVolt / SetCurr
 *                                                                         FDTD/engine.h:54-101
 *   Engine_HIP::CalcFastEnergy          Engine_Interface_FDTD::CalcFastEnergy
 *                                                                         FDTD/engine_interface_fdtd.cpp:302-347
 *   Engine_Interface_HIP probes         Engine_Interface_FDTD::CalcVoltageIntegral / GetRawField
 *                                                                         FDTD/engine_interface_fdtd.cpp:134-268
 *
 * The engine runs every extension of the operator on the device (excitation,
 * UPML, Mur, Lorentz / Drude / Debye, lumped RLC), so no host extension hook
 * exists.  Per-element field access goes through a host copy of the fields:
 * the first Get* after a step fetches it, Set* writes into it, and the next
 * IterateTS sends modified fields back first.
 *
 * Errors: where the reference returns a status (CalcECOperator: 0 or -1) the
 * same status is returned; everything else throws openems_hip::Error with the
 * library's message (the reference calls exit() on fatal set-up errors).
 */
#ifndef OPENEMS_HIP_HPP
#define OPENEMS_HIP_HPP

#include <algorithm>
#include <stdexcept>
#include <string>
#include <array>
#include <vector>

#include "ems_hip.h"
#include "ems_operator.h"

namespace openems_hip {

typedef float FDTD_FLOAT;

class Error : public std::runtime_error {
public:
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

inline void check_engine(int rc)
{
    if (rc < 0) throw Error(std::string("ems: ") + ems_last_error());
}

inline void check_operator(int rc)
{
    if (rc < 0) throw Error(std::string("emsop: ") + emsop_last_error());
}

class Engine_HIP;

// The operator of a box scene (materials, metals, lumped elements,
// excitations as axis-aligned boxes on a rectilinear mesh).
class Operator_HIP {
public:
    static Operator_HIP* New() { return new Operator_HIP(); }
    ~Operator_HIP() { emsop_destroy(h_); }
    Operator_HIP(const Operator_HIP&) = delete;
    Operator_HIP& operator=(const Operator_HIP&) = delete;

    // ---- scene (what SetGeometryCSX / SetupCSXGrid read from CSXCAD) ----------------
    void SetupCSXGrid(const std::vector<double> lines[3], double delta_unit)
    {
        check_operator(emsop_set_grid(h_, lines[0].data(), (unsigned)lines[0].size(), lines[1].data(),
                                      (unsigned)lines[1].size(), lines[2].data(), (unsigned)lines[2].size(),
                                      delta_unit));
    }
    void SetBackgroundMaterial(double eps_r, double mue_r, double kappa, double sigma)
    {
        check_operator(emsop_set_background(h_, eps_r, mue_r, kappa, sigma));
    }
    void AddMaterialBox(const double eps[3], const double mue[3], const double kappa[3], const double sigma[3],
                        int priority, const double start[3], const double stop[3])
    {
        check_operator(emsop_add_material(h_, eps, mue, kappa, sigma, priority, start, stop));
    }
    void AddMetalBox(int priority, const double start[3], const double stop[3])
    {
        check_operator(emsop_add_metal(h_, priority, start, stop));
    }
    void AddLumpedElement(int dir, double R, double C, double L, bool caps, bool parallel, int priority,
                          const double start[3], const double stop[3])
    {
        check_operator(emsop_add_lumped(h_, dir, R, C, L, caps, parallel, priority, start, stop));
    }
    // type 0 soft E, 1 hard E, 2 soft H, 3 hard H
    void AddExcitationBox(int type, const double val[3], double delay, int priority, const double start[3],
                          const double stop[3])
    {
        check_operator(emsop_add_excitation(h_, type, val, delay, priority, start, stop));
    }
    // wires (CSPrimCurve): points[i] in drawing units, as Operator::CalcPEC_Curves and
    // Operator_Ext_Excitation take them (FDTD/operator.cpp:2086-2118, operator_ext_excitation.cpp:235-287)
    void AddMetalCurve(const std::vector<std::array<double, 3>>& points)
    {
        check_operator(emsop_add_metal_curve(h_, (unsigned)points.size(), points.empty() ? nullptr : points[0].data()));
    }
    void AddExcitationCurve(int type, const double val[3], double delay, const std::vector<std::array<double, 3>>& points)
    {
        check_operator(emsop_add_excitation_curve(h_, type, val, delay, (unsigned)points.size(),
                                                  points.empty() ? nullptr : points[0].data()));
    }
    // shell_width 0: solid cylinder (CSPrimCylinder); > 0: cylindrical shell of mid radius `radius`
    void AddMetalCylinder(int priority, const double start[3], const double stop[3], double radius,
                          double shell_width = 0)
    {
        check_operator(emsop_add_metal_cylinder(h_, priority, start, stop, radius, shell_width));
    }
    void AddExcitationCylinder(int type, const double val[3], double delay, int priority, const double start[3],
                               const double stop[3], double radius, double shell_width = 0)
    {
        check_operator(emsop_add_excitation_cylinder(h_, type, val, delay, priority, start, stop, radius, shell_width));
    }
    // weight functions (x, y, z, rho, r, a) of the excitation primitive added last; an empty string means 1
    void SetExcitationWeight(const std::string& wx, const std::string& wy, const std::string& wz)
    {
        const char* w[3] = {wx.empty() ? nullptr : wx.c_str(), wy.empty() ? nullptr : wy.c_str(),
                            wz.empty() ? nullptr : wz.c_str()};
        check_operator(emsop_set_excitation_weight(h_, w));
    }
    // BC: 0 PEC, 1 PMC, 2 Mur, 3 PML, -1 none (openEMS::m_BC_type)
    void SetBoundaryCondition(const int BC[6], const unsigned pml_size[6], const double mur_vph[6])
    {
        check_operator(emsop_set_boundary(h_, BC, pml_size, mur_vph));
    }
    void SetupGaussianPulse(double f0, double fc) { check_operator(emsop_set_signal(h_, EMSOP_SIG_GAUSS, f0, fc)); }
    void SetupSinusoidal(double f0) { check_operator(emsop_set_signal(h_, EMSOP_SIG_SINUS, f0, 0)); }
    // Excitation::SetupDiracPulse / SetupStepExcite / SetupCustomExcite (FDTD/excitation.cpp:72-93)
    void SetupDiracPulse(double f_max) { check_operator(emsop_set_signal(h_, EMSOP_SIG_DIRAC, f_max, 0)); }
    void SetupStepExcite(double f_max) { check_operator(emsop_set_signal(h_, EMSOP_SIG_STEP, f_max, 0)); }
    void SetupCustomExcite(const std::string& function, double f0)
    {
        check_operator(emsop_set_custom_signal(h_, function.c_str(), f0));
    }
    void SetTimeStepMethod(int method) { ts_method_ = method; }
    void SetTimestepFactor(double factor) { ts_factor_ = factor; }

    // Operator::CalcECOperator: the operator and every extension table for a
    // run of nrts timesteps; 0 on success, -1 on failure (ems_last error text
    // in GetLastError)
    int CalcECOperator(unsigned nrts)
    {
        if (emsop_set_timestep(h_, ts_method_, 0.0, ts_factor_) < 0) return -1;
        return emsop_build(h_, nrts) < 0 ? -1 : 0;
    }
    static std::string GetLastError() { return emsop_last_error(); }

    // ---- queries (valid after CalcECOperator) ---------------------------------------
    double GetTimestep() const { return emsop_timestep(h_); }
    unsigned GetNumberOfLines(int ny) const
    {
        unsigned N[3];
        check_operator(emsop_num_lines(h_, N));
        return N[ny];
    }
    unsigned GetNumberCells() const { return GetNumberOfLines(0) * GetNumberOfLines(1) * GetNumberOfLines(2); }
    double GetDiscLine(int n, unsigned pos, bool dualMesh = false) const { return emsop_disc_line(h_, n, pos, dualMesh); }
    double GetEdgeLength(int n, const unsigned pos[3], bool dualMesh = false) const
    {
        return emsop_edge_length(h_, n, pos[n], dualMesh);
    }
    unsigned GetNyquistNum() const { return emsop_nyquist(h_); }
    FDTD_FLOAT GetVV(unsigned n, unsigned x, unsigned y, unsigned z) const { return coeff(0, n, x, y, z); }
    FDTD_FLOAT GetVI(unsigned n, unsigned x, unsigned y, unsigned z) const { return coeff(1, n, x, y, z); }
    FDTD_FLOAT GetII(unsigned n, unsigned x, unsigned y, unsigned z) const { return coeff(2, n, x, y, z); }
    FDTD_FLOAT GetIV(unsigned n, unsigned x, unsigned y, unsigned z) const { return coeff(3, n, x, y, z); }

    // Operator::CreateEngine: the device engine of this operator
    Engine_HIP* CreateEngine(int device = 0) const;

    emsop* handle() const { return h_; }

private:
    Operator_HIP() { check_operator(emsop_create(&h_)); }
    FDTD_FLOAT coeff(int which, unsigned n, unsigned x, unsigned y, unsigned z) const
    {
        const float* a = emsop_coeff(h_, which);
        if (!a) throw Error("operator not built");
        const unsigned Nx = GetNumberOfLines(0), Ny = GetNumberOfLines(1), Nz = GetNumberOfLines(2);
        return a[(((size_t)n * Nx + x) * Ny + y) * Nz + z];
    }
    emsop* h_ = nullptr;
    int ts_method_ = 3;
    double ts_factor_ = 1.0;
};

class Engine_HIP {
public:
    static Engine_HIP* New(const Operator_HIP* op, int device = 0)
    {
        Engine_HIP* e = new Engine_HIP(op, device);
        try {
            e->Init();
        } catch (...) {
            delete e;
            throw;
        }
        return e;
    }
    virtual ~Engine_HIP() { Reset(); }
    Engine_HIP(const Engine_HIP&) = delete;
    Engine_HIP& operator=(const Engine_HIP&) = delete;

    // Engine::Init: device fields and every extension of the operator
    virtual void Init()
    {
        if (h_) return;
        emsop* o = op_->handle();
        for (int n = 0; n < 3; ++n) N_[n] = op_->GetNumberOfLines(n);
        check_engine(ems_create(&h_, N_, emsop_coeff(o, 0), emsop_coeff(o, 1), emsop_coeff(o, 2),
                                emsop_coeff(o, 3), device_));
        // excitation (Operator_Ext_Excitation)
        const unsigned L = emsop_signal_length(o);
        std::vector<unsigned> vi[2];
        std::vector<unsigned short> vd[2];
        std::vector<float> va[2];
        std::vector<unsigned> vdl[2];
        for (int c = 0; c < 2; ++c) {
            const unsigned K = emsop_exc_count(o, c);
            vi[c].resize(3 * (size_t)K);
            vd[c].resize(K);
            va[c].resize(K);
            vdl[c].resize(K);
            if (K) check_operator(emsop_exc_get(o, c, vi[c].data(), vd[c].data(), va[c].data(), vdl[c].data()));
        }
        check_engine(ems_set_excitation(h_, emsop_signal_volt(o), emsop_signal_curr(o), L,
                                        emsop_signal_period_steps(o), (unsigned)vd[0].size(), vi[0].data(),
                                        vd[0].data(), va[0].data(), vdl[0].data(), (unsigned)vd[1].size(),
                                        vi[1].data(), vd[1].data(), va[1].data(), vdl[1].data()));
        // Mur faces (Operator_Ext_Mur_ABC), creation order
        for (unsigned k = 0; k < emsop_mur_count(o); ++k) {
            int ny;
            unsigned line, shift, start_ts;
            const float *cP, *cPP;
            check_operator(emsop_mur_get(o, k, &ny, &line, &shift, &start_ts, &cP, &cPP));
            check_engine(ems_add_mur(h_, ny, line, shift, cP, cPP, start_ts));
        }
        // UPML slabs (Operator_Ext_UPML)
        for (unsigned k = 0; k < emsop_upml_count(o); ++k) {
            unsigned st[3], num[3];
            const float* a[6];
            check_operator(emsop_upml_get(o, k, st, num, a));
            check_engine(ems_add_upml(h_, st, num, a[0], a[1], a[2], a[3], a[4], a[5]));
        }
        // Lorentz / Drude / Debye orders (Operator_Ext_LorentzMaterial)
        for (unsigned ord = 0; ord < emsop_lorentz_orders(o); ++ord) {
            unsigned cnt;
            int on[4];
            const unsigned* pos[3];
            const float* t[18];
            check_operator(emsop_lorentz_get(o, ord, &cnt, on, pos, t));
            if (!cnt || !(on[0] || on[1])) continue;
            std::vector<unsigned> P(3 * (size_t)cnt);
            std::vector<float> T[6];
            for (int n = 0; n < 3; ++n) std::copy(pos[n], pos[n] + cnt, P.begin() + (size_t)n * cnt);
            for (int k = 0; k < 6; ++k) {
                T[k].resize(3 * (size_t)cnt);
                for (int n = 0; n < 3; ++n) std::copy(t[3 * k + n], t[3 * k + n] + cnt, T[k].begin() + (size_t)n * cnt);
            }
            check_engine(ems_add_lorentz(h_, cnt, P.data(), T[0].data(), T[1].data(), on[2] ? T[4].data() : nullptr,
                                         T[2].data(), T[3].data(), on[3] ? T[5].data() : nullptr, on[0], on[1]));
        }
        // lumped RLC (Operator_Ext_LumpedRLC)
        if (const unsigned K = emsop_rlc_count(o)) {
            const unsigned* pos[3];
            const int* dir;
            const float* c[9];
            check_operator(emsop_rlc_get(o, pos, &dir, c));
            std::vector<unsigned> P(3 * (size_t)K);
            for (int n = 0; n < 3; ++n) std::copy(pos[n], pos[n] + K, P.begin() + (size_t)n * K);
            check_engine(ems_add_lumped_rlc(h_, K, P.data(), dir, c));
        }
        check_engine(ems_init(h_));
    }

    // Engine::Reset: releases the device state
    virtual void Reset()
    {
        if (h_) ems_destroy(h_);
        h_ = nullptr;
        row_valid_ = false;
    }

    // Engine::IterateTS
    virtual bool IterateTS(unsigned int iterTS)
    {
        check_engine(ems_iterate(h_, iterTS));
        row_valid_ = false;
        return true;
    }
    virtual unsigned int GetNumberOfTimesteps() { return ems_get_numts(h_); }

    // Engine::GetVolt / GetCurr: the x row holding the element is read from
    // the device once (ems_read_field) and kept until the next step or write
    FDTD_FLOAT GetVolt(unsigned n, unsigned x, unsigned y, unsigned z) { return row(false, n, y, z)[x]; }
    FDTD_FLOAT GetVolt(unsigned n, const unsigned pos[3]) { return GetVolt(n, pos[0], pos[1], pos[2]); }
    FDTD_FLOAT GetCurr(unsigned n, unsigned x, unsigned y, unsigned z) { return row(true, n, y, z)[x]; }
    FDTD_FLOAT GetCurr(unsigned n, const unsigned pos[3]) { return GetCurr(n, pos[0], pos[1], pos[2]); }
    // Engine::SetVolt / SetCurr: written through to the device (one element)
    void SetVolt(unsigned n, unsigned x, unsigned y, unsigned z, FDTD_FLOAT value) { set(false, n, x, y, z, value); }
    void SetVolt(unsigned n, const unsigned pos[3], FDTD_FLOAT value) { SetVolt(n, pos[0], pos[1], pos[2], value); }
    void SetCurr(unsigned n, unsigned x, unsigned y, unsigned z, FDTD_FLOAT value) { set(true, n, x, y, z, value); }
    void SetCurr(unsigned n, const unsigned pos[3], FDTD_FLOAT value) { SetCurr(n, pos[0], pos[1], pos[2], value); }

    // a box of voltages or currents, [3][nx][ny][nz] (comp -1) or [nx][ny][nz]
    void ReadBox(bool curr, int comp, const unsigned start[3], const unsigned stop[3], FDTD_FLOAT* out)
    {
        check_engine(ems_read_field(h_, curr ? 1 : 0, comp, start, stop, out));
    }
    void WriteBox(bool curr, int comp, const unsigned start[3], const unsigned stop[3], const FDTD_FLOAT* in)
    {
        check_engine(ems_write_field(h_, curr ? 1 : 0, comp, start, stop, in));
        row_valid_ = false;
    }

    // Engine_Interface_FDTD::CalcFastEnergy
    double CalcFastEnergy()
    {
        double e = 0;
        check_engine(ems_energy(h_, &e));
        return e;
    }

    ems_engine* handle() { return h_; }

protected:
    Engine_HIP(const Operator_HIP* op, int device) : op_(op), device_(device) {}

private:
    const std::vector<float>& row(bool curr, unsigned n, unsigned y, unsigned z)
    {
        if (!row_valid_ || row_key_[0] != (unsigned)curr || row_key_[1] != n || row_key_[2] != y || row_key_[3] != z) {
            const unsigned a[3] = {0, y, z}, b[3] = {N_[0] - 1, y, z};
            row_.resize(N_[0]);
            check_engine(ems_read_field(h_, curr ? 1 : 0, (int)n, a, b, row_.data()));
            row_key_[0] = curr;
            row_key_[1] = n;
            row_key_[2] = y;
            row_key_[3] = z;
            row_valid_ = true;
        }
        return row_;
    }
    void set(bool curr, unsigned n, unsigned x, unsigned y, unsigned z, float v)
    {
        const unsigned p[3] = {x, y, z};
        check_engine(ems_write_field(h_, curr ? 1 : 0, (int)n, p, p, &v));
        if (row_valid_ && row_key_[0] == (unsigned)curr && row_key_[1] == n && row_key_[2] == y && row_key_[3] == z)
            row_[x] = v;
    }

    const Operator_HIP* op_;
    int device_;
    ems_engine* h_ = nullptr;
    unsigned N_[3] = {0, 0, 0};
    std::vector<float> row_;
    unsigned row_key_[4] = {0, 0, 0, 0};
    bool row_valid_ = false;
};

inline Engine_HIP* Operator_HIP::CreateEngine(int device) const { return Engine_HIP::New(this, device); }

// Probes reduced on the device (Engine_Interface_FDTD, FDTD/engine_interface_fdtd.cpp):
// the voltage line integral of CalcVoltageIntegral (:206-232) and the point
// fields of GetRawField / GetRawDualField (:134-146, :263-268), sampled after
// every step that is a multiple of `interval` (and at step 0).  Register the
// probes before the first IterateTS.
class Engine_Interface_HIP {
public:
    Engine_Interface_HIP(Engine_HIP* eng, const Operator_HIP* op) : eng_(eng), op_(op) {}

    // line from start to stop (mesh indices, one axis differing; otherwise the
    // probe records zeros as the reference does); returns the probe id
    int AddVoltageProbe(const unsigned start[3], const unsigned stop[3], unsigned interval, unsigned capacity)
    {
        std::vector<unsigned> comp, pos[3];
        std::vector<int> sign;
        const int axes = (start[0] != stop[0]) + (start[1] != stop[1]) + (start[2] != stop[2]);
        for (int n = 0; n < 3 && axes == 1; ++n) {
            if (start[n] == stop[n]) continue;
            const bool up = start[n] < stop[n];
            unsigned p[3] = {up ? start[0] : stop[0], up ? start[1] : stop[1], up ? start[2] : stop[2]};
            for (; p[n] < (up ? stop[n] : start[n]); ++p[n]) {
                comp.push_back(n);
                for (int a = 0; a < 3; ++a) pos[a].push_back(p[a]);
                sign.push_back(up ? 1 : -1);
            }
        }
        std::vector<unsigned> P;
        for (int a = 0; a < 3; ++a) P.insert(P.end(), pos[a].begin(), pos[a].end());
        return add(EMS_PROBE_VOLTAGE, comp, P, sign, std::vector<double>(), interval, capacity);
    }
    // E (h_field = false) or H field at a mesh node: value / (dual) edge length
    int AddFieldProbe(bool h_field, const unsigned node[3], unsigned interval, unsigned capacity)
    {
        std::vector<unsigned> comp = {0, 1, 2}, P;
        for (int a = 0; a < 3; ++a)
            for (int n = 0; n < 3; ++n) P.push_back(node[a]);
        std::vector<double> div;
        for (int n = 0; n < 3; ++n) div.push_back(op_->GetEdgeLength(n, node, h_field));
        return add(h_field ? EMS_PROBE_HFIELD : EMS_PROBE_EFIELD, comp, P, std::vector<int>(3, 1), div, interval,
                   capacity);
    }
    // ProcessCurrent::CalcIntegral (Common/processcurrent.cpp:96-171): the
    // currents around the rectangle start..stop (snapped mesh indices, equal
    // along normDir), in the reference's order and orientation; returns the id
    int AddCurrentProbe(const unsigned start[3], const unsigned stop[3], int normDir, unsigned interval,
                        unsigned capacity)
    {
        if (normDir < 0 || normDir > 2 || start[normDir] != stop[normDir])
            throw Error("current probe: start and stop must lie in one plane normal to normDir");
        std::vector<unsigned> comp, pos[3];
        std::vector<int> sign;
        // the edges lo+1 .. hi along `axis` from the fixed point f, component c
        auto seg = [&](unsigned c, unsigned fx, unsigned fy, unsigned fz, int axis, unsigned lo, unsigned hi, int sg) {
            for (unsigned i = lo + 1; i <= hi; ++i) {
                unsigned p[3] = {fx, fy, fz};
                p[axis] = i;
                comp.push_back(c);
                for (int a = 0; a < 3; ++a) pos[a].push_back(p[a]);
                sign.push_back(sg);
            }
        };
        const unsigned* st = start;
        const unsigned* sp = stop;
        if (normDir == 0) {
            seg(1, sp[0], 0, st[2], 1, st[1], sp[1], 1);
            seg(2, sp[0], sp[1], 0, 2, st[2], sp[2], 1);
            seg(1, st[0], 0, sp[2], 1, st[1], sp[1], -1);
            seg(2, st[0], st[1], 0, 2, st[2], sp[2], -1);
        } else if (normDir == 1) {
            seg(2, st[0], st[1], 0, 2, st[2], sp[2], 1);
            seg(0, 0, sp[1], sp[2], 0, st[0], sp[0], 1);
            seg(2, sp[0], sp[1], 0, 2, st[2], sp[2], -1);
            seg(0, 0, st[1], st[2], 0, st[0], sp[0], -1);
        } else {
            seg(0, 0, st[1], st[2], 0, st[0], sp[0], 1);
            seg(1, sp[0], 0, st[2], 1, st[1], sp[1], 1);
            seg(0, 0, sp[1], sp[2], 0, st[0], sp[0], -1);
            seg(1, st[0], 0, sp[2], 1, st[1], sp[1], -1);
        }
        std::vector<unsigned> P;
        for (int a = 0; a < 3; ++a) P.insert(P.end(), pos[a].begin(), pos[a].end());
        return add(EMS_PROBE_CURRENT, comp, P, sign, std::vector<double>(), interval, capacity);
    }

    // Engine_Interface_Base::SetInterpolationType: 0 none, 1 node, 2 cell
    void SetInterpolationType(int type) { interp_ = type; }
    // Engine_Interface_FDTD::GetEField / GetHField at a mesh node (V/m, A/m)
    void GetEField(const unsigned pos[3], double out[3]) { field_at(0, pos, out); }
    void GetHField(const unsigned pos[3], double out[3]) { field_at(1, pos, out); }

    // samples so far: steps and values[count][3] (voltage probes use slot 0)
    std::vector<double> GetSamples(int id, std::vector<unsigned>* steps = nullptr)
    {
        unsigned count = 0;
        check_engine(ems_probe_results(eng_->handle(), id, &count, nullptr, nullptr));
        std::vector<unsigned> ts(count);
        std::vector<double> v(3 * (size_t)count);
        if (count) check_engine(ems_probe_results(eng_->handle(), id, &count, ts.data(), v.data()));
        if (steps) *steps = ts;
        return v;
    }

private:
    int add(int type, const std::vector<unsigned>& comp, const std::vector<unsigned>& pos, const std::vector<int>& sign,
            const std::vector<double>& div, unsigned interval, unsigned capacity)
    {
        ems_probe_desc d{};
        d.type = type;
        d.nelem = (unsigned)comp.size();
        d.comp = comp.data();
        d.pos = pos.data();
        d.sign = sign.data();
        d.divisor = div.empty() ? nullptr : div.data();
        d.interval = interval;
        d.fd_interval = 0;
        d.start_ts = 0;
        d.stop_ts = 0xFFFFFFFFu;
        d.capacity = capacity;
        int id = -1;
        check_engine(ems_add_probe(eng_->handle(), &d, &id));
        return id;
    }

    void field_at(int field, const unsigned pos[3], double out[3])
    {
        std::vector<double>* el = field ? dual_ : prim_;
        if (el[0].empty()) {
            for (int n = 0; n < 3; ++n)
                for (unsigned i = 0; i < op_->GetNumberOfLines(n); ++i) {
                    unsigned p[3] = {0, 0, 0};
                    p[n] = i;
                    el[n].push_back(op_->GetEdgeLength(n, p, field == 1));
                }
        }
        const double* e[3] = {el[0].data(), el[1].data(), el[2].data()};
        check_engine(ems_get_field_at(eng_->handle(), field, interp_, pos, e, out));
    }

    Engine_HIP* eng_;
    const Operator_HIP* op_;
    int interp_ = 0;                       // no interpolation, the reference's default (engine_interface_base.cpp:24)
    std::vector<double> prim_[3], dual_[3];
};

}  // namespace openems_hip

#endif /* OPENEMS_HIP_HPP */
