// Runs a small scene through the C++ host classes of include/openems_hip.hpp
// (Operator_HIP -> Engine_HIP), with per-element Get/Set between iterations,
// and prints the values it reads as one JSON object.  The GPU test runs the
// same scene and calls through the Python host and compares bit for bit.
//
// Scene (drawing unit 1 mm, 28 lines per axis, uniform 1 mm): 6-cell UPML on
// the x faces, PEC elsewhere; a dielectric box (eps_r 3, kappa 0.02) of
// priority 1 from (8, 8, 8) to (18, 16, 20); a metal box of priority 2 from
// (12, 20, 4) to (14, 22, 24); a 50 Ohm parallel lumped port along z at
// (14, 14, 10)-(14, 14, 12) with caps and a soft E_z source on the same box;
// a Drude box (plasma 12 GHz, relaxation 30 ps) from (19, 9, 9) to (22, 14, 18);
// a metal wire (curve) and a metal cylinder (radius 1.6, axis (20, 18, 3) to
// (24, 24, 9)); a soft E_x source on a cylinder (radius 1.2, axis (5, 21, 21)
// to (10, 21, 21)) weighted 1 + 0.1 x;
// Gaussian pulse f0 = 0, fc = 15 GHz.  Probes: a voltage line across the
// port, E and H points, a current loop around the port; after the last step
// the E and H fields at three nodes with each interpolation type.
#include <cstdio>
#include <array>
#include <vector>

#include "openems_hip.hpp"

using namespace openems_hip;

static void hexf(const char* key, float v, bool comma = true)
{
    std::printf("\"%s\": \"%a\"%s", key, (double)v, comma ? ", " : "");
}

int main()
{
    Operator_HIP* op = Operator_HIP::New();
    std::vector<double> lines[3];
    for (int n = 0; n < 3; ++n)
        for (int i = 0; i < 28; ++i) lines[n].push_back(i);
    op->SetupCSXGrid(lines, 1e-3);
    op->SetBackgroundMaterial(1, 1, 0, 0);
    const double eps[3] = {3, 3, 3}, mue[3] = {1, 1, 1}, kap[3] = {0.02, 0.02, 0.02}, sig[3] = {0, 0, 0};
    const double d0[3] = {8, 8, 8}, d1[3] = {18, 16, 20};
    op->AddMaterialBox(eps, mue, kap, sig, 1, d0, d1);
    const double one[3] = {1, 1, 1}, zero[3] = {0, 0, 0};
    const double r0[3] = {19, 9, 9}, r1[3] = {22, 14, 18};
    op->AddMaterialBox(one, one, zero, zero, 3, r0, r1);
    double disp[18] = {0};
    for (int n = 0; n < 3; ++n) {
        disp[0 * 3 + n] = 12e9;     // eps plasma frequency
        disp[1 * 3 + n] = 30e-12;   // eps relaxation time
    }
    if (emsop_set_dispersion(op->handle(), 1, disp) < 0) return 2;
    const double m0[3] = {12, 20, 4}, m1[3] = {14, 22, 24};
    op->AddMetalBox(2, m0, m1);
    op->AddMetalCurve({{{4.0, 4.0, 6.0}, {4.0, 4.0, 12.0}, {7.3, 5.6, 15.2}}});
    const double c0[3] = {20, 18, 3}, c1[3] = {24, 24, 9};
    op->AddMetalCylinder(1, c0, c1, 1.6);
    const double p0[3] = {14, 14, 10}, p1[3] = {14, 14, 12};
    op->AddLumpedElement(2, 50.0, 0.0, 0.0, true, true, 0, p0, p1);
    const double ez[3] = {0, 0, 1};
    op->AddExcitationBox(0, ez, 0.0, 0, p0, p1);
    const double ex[3] = {1, 0, 0}, x0[3] = {5, 21, 21}, x1[3] = {10, 21, 21};
    op->AddExcitationCylinder(0, ex, 0.0, 0, x0, x1, 1.2);
    op->SetExcitationWeight("1+0.1*x", "", "");
    const int bc[6] = {3, 3, 0, 0, 0, 0};
    const unsigned pml[6] = {6, 6, 6, 6, 6, 6};
    const double vph[6] = {0, 0, 0, 0, 0, 0};
    op->SetBoundaryCondition(bc, pml, vph);
    op->SetupGaussianPulse(0.0, 15e9);
    if (op->CalcECOperator(400) != 0) {
        std::fprintf(stderr, "CalcECOperator failed: %s\n", Operator_HIP::GetLastError().c_str());
        return 1;
    }
    try {
        Engine_HIP* eng = op->CreateEngine();
        eng->Init();
        Engine_Interface_HIP iface(eng, op);
        const unsigned every = op->GetNyquistNum() / 4;
        const unsigned ua[3] = {14, 14, 10}, ub[3] = {14, 14, 12}, en[3] = {20, 11, 12}, hn[3] = {10, 12, 14};
        const int pu = iface.AddVoltageProbe(ua, ub, every, 512);
        const int pe = iface.AddFieldProbe(false, en, every, 512);
        const int ph = iface.AddFieldProbe(true, hn, every, 512);
        const unsigned ia[3] = {13, 13, 11}, ib[3] = {14, 14, 11};
        const int pi = iface.AddCurrentProbe(ia, ib, 2, every, 512);
        eng->IterateTS(120);
        std::printf("{");
        std::printf("\"dT\": \"%a\", ", op->GetTimestep());
        std::printf("\"numTS_a\": %u, ", eng->GetNumberOfTimesteps());
        hexf("vi_a", op->GetVI(2, 14, 14, 10));
        hexf("v_a", eng->GetVolt(2, 14, 14, 11));
        hexf("v_b", eng->GetVolt(0, 20, 11, 12));
        hexf("i_a", eng->GetCurr(1, 10, 12, 14));
        std::printf("\"energy_a\": \"%a\", ", eng->CalcFastEnergy());
        eng->SetVolt(0, 5, 5, 5, 0.25f);
        const unsigned q[3] = {7, 9, 11};
        eng->SetCurr(1, q, -0.5f);
        eng->IterateTS(80);
        std::printf("\"numTS_b\": %u, ", eng->GetNumberOfTimesteps());
        hexf("v_c", eng->GetVolt(2, 14, 14, 11));
        hexf("v_d", eng->GetVolt(0, 5, 5, 5));
        hexf("i_b", eng->GetCurr(1, q));
        hexf("v_e", eng->GetVolt(1, 21, 12, 13));
        // interpolated point fields (Engine_Interface_FDTD::GetEField / GetHField)
        std::printf("\"fields\": [");
        const unsigned fpos[3][3] = {{20, 11, 12}, {10, 12, 14}, {27, 0, 5}};
        bool first = true;
        for (int it = 0; it < 3; ++it) {
            iface.SetInterpolationType(it);
            for (const auto& fp : fpos) {
                double e[3], hh[3];
                iface.GetEField(fp, e);
                iface.GetHField(fp, hh);
                for (double v : e) std::printf("%s\"%a\"", first ? "" : ", ", v), first = false;
                for (double v : hh) std::printf(", \"%a\"", v);
            }
        }
        std::printf("], ");
        std::vector<unsigned> ts;
        const std::vector<double> su = iface.GetSamples(pu, &ts), se = iface.GetSamples(pe), sh = iface.GetSamples(ph);
        const std::vector<double> si = iface.GetSamples(pi);
        std::printf("\"probe_every\": %u, \"probe_count\": %u, ", every, (unsigned)ts.size());
        std::printf("\"probe_u\": [");
        for (size_t k = 0; k < ts.size(); ++k) std::printf("%s\"%a\"", k ? ", " : "", su[3 * k]);
        std::printf("], \"probe_e\": [");
        for (size_t k = 0; k < 3 * ts.size(); ++k) std::printf("%s\"%a\"", k ? ", " : "", se[k]);
        std::printf("], \"probe_h\": [");
        for (size_t k = 0; k < 3 * ts.size(); ++k) std::printf("%s\"%a\"", k ? ", " : "", sh[k]);
        std::printf("], \"probe_i\": [");
        for (size_t k = 0; k < ts.size(); ++k) std::printf("%s\"%a\"", k ? ", " : "", si[3 * k]);
        std::printf("], ");
        std::printf("\"energy_b\": \"%a\"}\n", eng->CalcFastEnergy());
        delete eng;
    } catch (const Error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        delete op;
        return 1;
    }
    delete op;
    return 0;
}
