// sanitize_check.cpp -- TEST INFRASTRUCTURE: runs the oracle (oracle.cpp) built with the
// host address and undefined-behaviour sanitizers over small cases of every operator,
// precision, boundary option and strip head, so an out-of-bounds index or an undefined
// operation in the restatement fails the CPU test suite (tests/test_oracle_sanitizers.py).
// Exit status 0 when every run succeeded and produced finite outputs.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "oracle.h"

namespace {

unsigned g_seed = 12345u;
float urand() {   // deterministic LCG in [0, 1)
    g_seed = g_seed * 1664525u + 1013904223u;
    return (float)((g_seed >> 8) & 0xffffff) / 16777216.0f;
}

struct Model {
    std::vector<float> vel, rho, eps, del, th, ph;
};

Model make_model(int n) {
    Model m;
    for (int i = 0; i < n; ++i) {
        const float v = 2000.0f + 1000.0f * urand();
        m.vel.push_back(v);
        m.rho.push_back(310.0f * std::pow(v, 0.25f));
        const float e = 0.2f * urand();
        m.eps.push_back(e);
        m.del.push_back(std::fmin(0.1f * urand(), e));
        m.th.push_back(0.7f * urand());
        m.ph.push_back(6.2f * urand());
    }
    return m;
}

int run_case(int op, int prec, int nb, int mute, int snaps, int sv, int wtype) {
    OracleShot s{};
    s.op = op;
    s.nx = 14;
    s.ny = 12;
    s.nz = 13;
    s.dx = s.dy = s.dz = 10.0f;
    s.nb = nb;
    s.sx = op == ORC_OP_2ND ? 6 : 7;
    s.sy = op == ORC_OP_2ND ? 5 : 6;
    s.sz = op == ORC_OP_2ND ? 2 : 3;
    s.nt = 40;
    s.dt = 1e-3f;
    s.fm = 25.0f;
    s.wtype = wtype;
    s.pfac = op == ORC_OP_ISO ? 1e4f : 10.0f;
    s.rec_z = op == ORC_OP_2ND ? 1 : 0;
    s.sv = sv;
    s.snap_interval = snaps ? 20 : 0;
    s.snap_skip0 = 1;
    s.mute = mute;
    s.v_mute = mute ? 2000.0f : 0.0f;
    s.mute_tt = 60;
    s.nthreads = 1;
    const int n = s.nx * s.ny * s.nz;
    const Model m = make_model(n);
    const size_t el = prec == 64 ? 8 : 4;
    std::vector<unsigned char> shot((size_t)s.nx * s.ny * s.nt * el);
    const int nsnap = snaps ? 2 : 0;
    std::vector<unsigned char> sn((size_t)(nsnap ? nsnap : 1) * n * el);
    double secs = 0.0;
    const bool vti = op == ORC_OP_VTI || op == ORC_OP_TTI;
    const int rc = oracle_run(prec, &s, m.vel.data(), op == ORC_OP_ISO || op == ORC_OP_VTI ? m.rho.data() : nullptr,
                              vti ? m.eps.data() : nullptr, vti ? m.del.data() : nullptr,
                              op == ORC_OP_TTI ? m.th.data() : nullptr, op == ORC_OP_TTI ? m.ph.data() : nullptr,
                              shot.data(), nsnap ? sn.data() : nullptr, nsnap, &secs);
    if (rc != (snaps ? 1 : 0)) {   // oracle_run returns the snapshot count (it = 20 of 40)
        std::printf("op %d prec %d nb %d: oracle_run returned %d\n", op, prec, nb, rc);
        return 1;
    }
    double sum = 0.0;
    for (size_t i = 0; i < shot.size() / el; ++i)
        sum += prec == 64 ? ((const double*)shot.data())[i] : ((const float*)shot.data())[i];
    if (!std::isfinite(sum)) {
        std::printf("op %d prec %d nb %d: non-finite record\n", op, prec, nb);
        return 1;
    }
    std::printf("op %d prec %d nb %d mute %d snaps %d sv %d wtype %d: ok (record sum %.3e)\n", op, prec, nb, mute,
                snaps, sv, wtype, sum);
    return 0;
}

}  // namespace

int main() {
    int bad = 0;
    for (int prec : {32, 64}) {
        bad += run_case(ORC_OP_2ND, prec, 8, 0, 1, 0, 1);
        bad += run_case(ORC_OP_VTI, prec, 8, 1, 1, 1, 1);
        bad += run_case(ORC_OP_TTI, prec, 8, 1, 1, 1, 1);
    }
    for (int w = 1; w <= 6; ++w) bad += run_case(ORC_OP_ISO, 32, 8, w == 1, w == 2, 0, w);
    // the learned strip (VTI, nb 10) under each head, with a seeded network
    const int np = 64 * 16 + 64 + 64 * 64 + 64 + 2 * 64 + 2;
    std::vector<float> net(np);
    for (auto& v : net) v = 0.2f * (urand() - 0.5f);
    for (int head = 0; head <= 2; ++head) {
        oracle_set_strip_net(net.data(), np);
        oracle_set_strip_head(head);
        oracle_set_strip_half(head != 1);
        oracle_set_strip_period(head == 2 ? 3 : 1);
        bad += run_case(ORC_OP_VTI, 32, 10, 0, 1, 0, 1);
    }
    oracle_set_strip_net(nullptr, 0);
    if (bad) std::printf("sanitize check: %d failing cases\n", bad);
    else std::printf("sanitize check: all cases ok\n");
    return bad ? 1 : 0;
}
