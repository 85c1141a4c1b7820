// host_check.cpp -- the host-only set-up logic of libsfem_hip.so
// (solidfem_amd/csrc/sfem_host.cpp) built for the CPU alone and run under
// AddressSanitizer + UndefinedBehaviorSanitizer by tests/test_host_sanitized_cpu.py.
// Checks the invariants the device kernels rely on, on structured, jittered
// and irregular meshes (nodes no element references, a single element,
// several blocks, varying moduli), the CSR tables, and the multigrid level
// set-up; exits non-zero on the first violation.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <set>
#include <vector>

#include "sfem_host.hpp"

using namespace sfem;

static int failures = 0;
#define CHECK(cond)                                                                  \
    do {                                                                             \
        if (!(cond)) {                                                               \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            failures++;                                                              \
            return;                                                                  \
        }                                                                            \
    } while (0)

struct Mesh {
    std::vector<int32_t> tets;
    std::vector<double> X;
    std::vector<uint8_t> fixed;
    int64_t ne() const { return (int64_t)tets.size() / 4; }
    int64_t nn() const { return (int64_t)X.size() / 3; }
};

// a hex grid split into 6 tets per cell (positive volumes), x = 0 face fixed
static Mesh beam(int a, int b, int c, double jitter, unsigned seed)
{
    Mesh m;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> U(-1, 1);
    auto id = [&](int i, int j, int k) { return (i * (b + 1) + j) * (c + 1) + k; };
    for (int i = 0; i <= a; i++)
        for (int j = 0; j <= b; j++)
            for (int k = 0; k <= c; k++) {
                const bool inner = i > 0 && i < a && j > 0 && j < b && k > 0 && k < c;
                m.X.push_back(i + (inner ? jitter * U(rng) : 0));
                m.X.push_back(j + (inner ? jitter * U(rng) : 0));
                m.X.push_back(k + (inner ? jitter * U(rng) : 0));
                m.fixed.push_back(i == 0);
            }
    for (int i = 0; i < a; i++)
        for (int j = 0; j < b; j++)
            for (int k = 0; k < c; k++) {
                const int v[8] = {id(i, j, k), id(i + 1, j, k), id(i, j + 1, k), id(i + 1, j + 1, k),
                                  id(i, j, k + 1), id(i + 1, j, k + 1), id(i, j + 1, k + 1), id(i + 1, j + 1, k + 1)};
                // the 6 tets around the diagonal v0 - v7
                const int t[6][4] = {{0, 1, 3, 7}, {0, 3, 2, 7}, {0, 2, 6, 7}, {0, 6, 4, 7}, {0, 4, 5, 7}, {0, 5, 1, 7}};
                for (auto& q : t) {
                    int n[4] = {v[q[0]], v[q[1]], v[q[2]], v[q[3]]};
                    // orient positively
                    double D[9];
                    for (int kk = 1; kk < 4; kk++)
                        for (int r = 0; r < 3; r++) D[3 * r + kk - 1] = m.X[3 * n[kk] + r] - m.X[3 * n[0] + r];
                    const double det = D[0] * (D[4] * D[8] - D[5] * D[7]) - D[1] * (D[3] * D[8] - D[5] * D[6]) +
                                       D[2] * (D[3] * D[7] - D[4] * D[6]);
                    if (det < 0) std::swap(n[0], n[1]);
                    m.tets.insert(m.tets.end(), n, n + 4);
                }
            }
    return m;
}

static void check_mesh(const Mesh& m, bool vary_moduli, const char* name)
{
    const int64_t ne = m.ne(), nn = m.nn();
    std::vector<double> E(ne, 1e6), nu(ne, 0.45);
    if (vary_moduli)
        for (int64_t e = 0; e < ne; e++) { E[e] = 1e5 * (1 + e % 7); nu[e] = 0.3 + 0.01 * (e % 13); }
    ElementGeometry g;
    std::string err;
    CHECK(element_geometry(m.tets.data(), ne, m.X.data(), nn, m.fixed.data(), 1e6, 0.45, vary_moduli ? E.data() : nullptr,
                           vary_moduli ? nu.data() : nullptr, 1070.0, &g, &err));
    // Dm^-1 inverts Dm; volumes positive; masses sum to rho V
    double vtot = 0, mtot = 0, mfree_expect = 0;
    std::vector<int> used(nn, 0);
    for (int64_t e = 0; e < ne; e++) {
        const int32_t* t = &m.tets[4 * e];
        double D[9];
        for (int k = 1; k < 4; k++)
            for (int r = 0; r < 3; r++) D[3 * r + k - 1] = m.X[3 * t[k] + r] - m.X[3 * t[0] + r];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double s = 0;
                for (int k = 0; k < 3; k++) s += D[3 * i + k] * g.dminv[9 * e + 3 * k + j];
                CHECK(std::fabs(s - (i == j)) < 1e-10);
            }
        CHECK(g.vol[e] > 0);
        vtot += g.vol[e];
        for (int k = 0; k < 4; k++) used[t[k]] = 1;
    }
    for (int64_t i = 0; i < nn; i++) {
        CHECK(g.held[i] == (m.fixed[i] || !used[i]));
        mtot += g.mass[i];
        CHECK((g.mass[i] > 0) == (g.held[i] == 0));
    }
    (void)mfree_expect;
    CHECK(mtot <= 1070.0 * vtot * (1 + 1e-12));

    for (int BE : {256, 64}) {
        const Partition P = make_partition(m.tets.data(), ne, m.X.data(), nn, g.held.data(), g.mu, g.lam, g.dminv, g.vol, BE);
        CHECK(P.nblk == (int)((ne + BE - 1) / BE) && P.ne_pad == (int64_t)P.nblk * BE);
        // every element exactly once, in block order, its local ids naming its nodes
        std::vector<int> seen(ne, 0);
        std::map<int, std::set<int>> node_blocks;
        for (int b = 0; b < P.nblk; b++) {
            const int n0 = P.bnode_ptr[b], nb = P.bnode_ptr[b + 1] - n0;
            CHECK(nb > 0 && nb <= P.nmax && nb < 65536);
            // the block's rows name distinct nodes (in an LDS bank-aware order, not sorted)
            std::vector<int> rows_sorted(nb);
            for (int r = 0; r < nb; r++) rows_sorted[r] = (int)(P.bnodes[n0 + r] & F_NODE_MASK);
            std::sort(rows_sorted.begin(), rows_sorted.end());
            for (int r = 0; r + 1 < nb; r++) CHECK(rows_sorted[r] < rows_sorted[r + 1]);
            // incidence offsets: this block's lptr segment is at n0 + b
            const uint16_t* lp = &P.lptr[n0 + b];
            CHECK(lp[0] == 0);
            std::vector<int> slot_used(lp[nb], 0);
            for (int64_t s = (int64_t)b * BE; s < std::min<int64_t>(ne, (int64_t)(b + 1) * BE); s++) {
                const int e = P.eorig[s];
                CHECK(e >= 0 && e < ne);
                seen[e]++;
                const uint32_t* li = &P.lidx[2 * s];
                const uint32_t* lq = &P.lperm[2 * s];
                const int l[4] = {(int)(li[0] & 0xffff), (int)(li[0] >> 16), (int)(li[1] & 0xffff), (int)(li[1] >> 16)};
                const int q[4] = {(int)(lq[0] & 0xffff), (int)(lq[0] >> 16), (int)(lq[1] & 0xffff), (int)(lq[1] >> 16)};
                for (int k = 0; k < 4; k++) {
                    CHECK(l[k] < nb);
                    const uint32_t row = P.bnodes[n0 + l[k]];
                    CHECK((int)(row & F_NODE_MASK) == m.tets[4 * e + k]);
                    CHECK(((row & F_FIXED) != 0) == (g.held[row & F_NODE_MASK] != 0));
                    // the corner's incidence slot lies in its node's range, once
                    CHECK(q[k] >= lp[l[k]] && q[k] < lp[l[k] + 1]);
                    CHECK(slot_used[q[k]]++ == 0);
                    node_blocks[row & F_NODE_MASK].insert(b);
                }
                for (int k = 0; k < 9; k++) CHECK(P.dm64[(size_t)k * P.ne_pad + s] == g.dminv[9 * e + k]);
                CHECK(P.cf64[s] == g.vol[e] * g.mu[e]);
                if (!P.uniform) CHECK(P.mat32[P.ne_pad + s] == (float)g.lam[e]);
            }
            for (int u : slot_used) CHECK(u == 1);
        }
        for (int64_t e = 0; e < ne; e++) CHECK(seen[e] == 1);
        CHECK(P.uniform == !vary_moduli);
        // shared nodes: one slot per block row, contiguous per node, in block order
        int nslots = 0;
        for (size_t i = 0; i < P.sh_node.size(); i++) {
            const int node = (int)(P.sh_node[i] & F_NODE_MASK);
            CHECK(node_blocks[node].size() > 1);
            CHECK(P.sh_ptr[i + 1] - P.sh_ptr[i] == (int)node_blocks[node].size());
            // a held node's row of K d is 0: the CG update reads no slots for it
            if (g.held[node]) CHECK(P.node_q[2 * node] == -2);
            else CHECK(P.node_q[2 * node] == P.sh_ptr[i] && P.node_q[2 * node + 1] == P.sh_ptr[i + 1]);
            nslots += P.sh_ptr[i + 1] - P.sh_ptr[i];
        }
        CHECK(nslots == P.sh_ptr.back());
        std::vector<int> slot_rows(nslots, 0);
        for (size_t r = 0; r < P.bnodes.size(); r++) {
            const int node = (int)(P.bnodes[r] & F_NODE_MASK);
            const bool shared = node_blocks[node].size() > 1;
            CHECK(((P.bnodes[r] & F_SHARED) != 0) == shared);
            if (shared) {
                if (!g.held[node]) CHECK(P.bslot[r] >= P.node_q[2 * node] && P.bslot[r] < P.node_q[2 * node + 1]);
                slot_rows[P.bslot[r]]++;
            } else {
                CHECK(P.bslot[r] == -1);
            }
        }
        for (int u : slot_rows) CHECK(u == 1);
        // one owner row per referenced node: the row of its only block, or of
        // the first of its blocks (the one holding its first partial slot)
        std::vector<int> owners(nn, 0), first_block(nn, 1 << 30);
        for (int b = 0; b < P.nblk; b++)
            for (int r = P.bnode_ptr[b]; r < P.bnode_ptr[b + 1]; r++) {
                const int node = (int)(P.bnodes[r] & F_NODE_MASK);
                first_block[node] = std::min(first_block[node], b);
            }
        for (int b = 0; b < P.nblk; b++)
            for (int r = P.bnode_ptr[b]; r < P.bnode_ptr[b + 1]; r++) {
                const int node = (int)(P.bnodes[r] & F_NODE_MASK);
                if (P.bnodes[r] & F_OWNER) {
                    owners[node]++;
                    CHECK(b == first_block[node]);
                }
            }
        for (int64_t i = 0; i < nn; i++) CHECK(owners[i] == (used[i] ? 1 : 0));
        for (int64_t i = 0; i < nn; i++) {
            if (g.held[i] || !used[i]) CHECK(P.node_q[2 * i] == -2);
            else if (node_blocks[i].size() == 1) CHECK(P.node_q[2 * i] == -1);
        }
        CHECK(P.node_q[2 * nn] == 0 && P.node_q[2 * nn + 1] == 0);
        // rest coordinates relative to each block's box centre reproduce the edges
        for (int b = 0; b < P.nblk; b++)
            for (int r = P.bnode_ptr[b]; r + 1 < P.bnode_ptr[b + 1]; r++) {
                const int a0 = (int)(P.bnodes[r] & F_NODE_MASK), a1 = (int)(P.bnodes[r + 1] & F_NODE_MASK);
                for (int k = 0; k < 3; k++) {
                    const double edge = m.X[3 * a1 + k] - m.X[3 * a0 + k];
                    CHECK(std::fabs((double)P.rel32[3 * (r + 1) + k] - (double)P.rel32[3 * r + k] - edge) <=
                          1e-6 * (1 + std::fabs(edge)));
                }
            }
    }
    std::printf("%s: %lld elements, %lld nodes ok\n", name, (long long)ne, (long long)nn);
}

static void check_csr_helpers()
{
    // pressure triangles: every (node, triangle) incidence once, triangle order
    const int32_t tris[] = {0, 1, 2, 2, 1, 3, 3, 4, 0};
    std::vector<int> ptr, idx;
    node_triangle_csr(tris, 3, 6, &ptr, &idx);
    CHECK(ptr.size() == 7 && ptr.back() == 9 && ptr[5] == ptr[6]);
    for (int n = 0; n < 6; n++)
        for (int k = ptr[n]; k < ptr[n + 1]; k++) {
            const int t = idx[k];
            CHECK(tris[3 * t] == n || tris[3 * t + 1] == n || tris[3 * t + 2] == n);
            if (k > ptr[n]) CHECK(idx[k - 1] < t);
        }
    // restriction = transpose of the prolongation, rows increasing per column
    const int64_t pp[] = {0, 2, 3, 5, 5};
    const int32_t pc[] = {0, 1, 1, 0, 2};
    const double pv[] = {0.5, 0.5, 1.0, 0.25, 0.75};
    std::vector<int> tp, tr;
    std::vector<float> tv;
    transpose_prolongation(pp, pc, pv, 4, 3, &tp, &tr, &tv);
    CHECK(tp.size() == 4 && tp[0] == 0 && tp[3] == 5);
    CHECK(tr[tp[0]] == 0 && tr[tp[0] + 1] == 2 && tv[tp[0] + 1] == 0.25f);
    CHECK(tr[tp[1]] == 0 && tr[tp[1] + 1] == 1 && tr[tp[2]] == 2 && tv[tp[2]] == 0.75f);
    // cable attachments: each (cable node, corner) once, nodes increasing
    const std::vector<int> node4 = {5, 2, 7, 1, 2, 3, 4, 5};
    std::vector<int> touched, tptr;
    std::vector<std::pair<int, int>> tent;
    cable_attachments(node4, 8, &touched, &tptr, &tent);
    CHECK(touched == (std::vector<int>{1, 2, 3, 4, 5, 7}));
    CHECK(tptr.size() == touched.size() + 1 && tptr.back() == 8);
    for (size_t t = 0; t < touched.size(); t++)
        for (int k = tptr[t]; k < tptr[t + 1]; k++) CHECK(node4[tent[k].second] == touched[t] && tent[k].second / 4 == tent[k].first);
    std::printf("csr helpers ok\n");
}

// the directly solved multigrid level's element data and incidence lists,
// and the validation of a level's transfer arguments
static void check_multigrid_setup()
{
    const Mesh m = beam(3, 2, 2, 0.15, 11);
    std::vector<double> ye(m.ne()), pe(m.ne());
    for (int64_t e = 0; e < m.ne(); e++) {
        ye[e] = 1e5 * (1 + e % 7);
        pe[e] = 0.3 + 0.01 * (e % 5);
    }
    const DenseLevelData d = dense_level_data(m.tets.data(), m.ne(), m.X.data(), m.nn(), 1e6, 0.45, ye.data(), pe.data());
    CHECK(d.tets4 == m.tets);
    CHECK((int64_t)d.inc_ptr.size() == m.nn() + 1 && d.inc_ptr.back() == 4 * m.ne());
    std::vector<int> seen(4 * m.ne(), 0);
    for (int64_t n = 0; n < m.nn(); n++)
        for (int k = d.inc_ptr[n]; k < d.inc_ptr[n + 1]; k++) {
            const int e = d.inc[2 * k], a = d.inc[2 * k + 1];
            CHECK(e >= 0 && e < m.ne() && a >= 0 && a < 4 && m.tets[4 * e + a] == n);
            if (k > d.inc_ptr[n]) CHECK(d.inc[2 * (k - 1)] <= e);    // element order
            seen[4 * e + a]++;
        }
    for (int u : seen) CHECK(u == 1);
    for (int64_t e = 0; e < m.ne(); e++) {
        const int32_t* t = m.tets.data() + 4 * e;
        double D[9];
        for (int k = 1; k < 4; k++)
            for (int r = 0; r < 3; r++) D[3 * r + (k - 1)] = m.X[3 * t[k] + r] - m.X[3 * t[0] + r];
        // Dm^-1 Dm = I (FP32 storage)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double v = 0;
                for (int k = 0; k < 3; k++) v += d.dminv[9 * e + 3 * i + k] * D[3 * k + j];
                CHECK(std::fabs(v - (i == j)) < 1e-5);
            }
        const double det = D[0] * (D[4] * D[8] - D[5] * D[7]) - D[1] * (D[3] * D[8] - D[5] * D[6]) +
                           D[2] * (D[3] * D[7] - D[4] * D[6]);
        const double mu = 0.5 * ye[e] / (1 + pe[e]), lam = ye[e] * pe[e] / (1 + pe[e]) / (1 - 2 * pe[e]);
        CHECK(std::fabs(d.cvol[2 * e] - det / 6 * mu) <= 1e-6 * det / 6 * mu);
        CHECK(std::fabs(d.cvol[2 * e + 1] - det / 6 * lam) <= 1e-6 * det / 6 * lam);
    }
    // transfer validation: a valid 5 -> 3 injection and prolongation, then each defect
    const int32_t inj[] = {0, 2, 4};
    const int64_t pp[] = {0, 1, 3, 4, 6, 7};
    const int32_t pc[] = {0, 0, 1, 1, 1, 2, 2};
    std::string err;
    CHECK(check_level_transfer(5, 3, inj, pp, pc, &err));
    CHECK(!check_level_transfer(3, 3, inj, pp, pc, &err) && err.find("fewer") != std::string::npos);
    const int32_t bad_inj[] = {0, 2, 5};
    CHECK(!check_level_transfer(5, 3, bad_inj, pp, pc, &err) && err.find("injection") != std::string::npos);
    const int64_t pp_start[] = {1, 1, 3, 4, 6, 7};
    CHECK(!check_level_transfer(5, 3, inj, pp_start, pc, &err) && err.find("p_ptr[0]") != std::string::npos);
    const int64_t pp_down[] = {0, 3, 2, 4, 6, 7};
    CHECK(!check_level_transfer(5, 3, inj, pp_down, pc, &err) && err.find("monotone") != std::string::npos);
    const int32_t pc_bad[] = {0, 0, 1, 3, 1, 2, 2};
    CHECK(!check_level_transfer(5, 3, inj, pp, pc_bad, &err) && err.find("column") != std::string::npos);
    std::printf("multigrid set-up ok\n");
}

static void check_errors()
{
    Mesh m = beam(1, 1, 1, 0.0, 1);
    ElementGeometry g;
    std::string err;
    m.tets[5] = (int32_t)m.nn();   // out of range
    CHECK(!element_geometry(m.tets.data(), m.ne(), m.X.data(), m.nn(), m.fixed.data(), 1e6, 0.3, nullptr, nullptr, 1000, &g,
                            &err) && err.find("range") != std::string::npos);
    m = beam(1, 1, 1, 0.0, 1);
    m.tets[3] = m.tets[2];          // degenerate
    CHECK(!element_geometry(m.tets.data(), m.ne(), m.X.data(), m.nn(), m.fixed.data(), 1e6, 0.3, nullptr, nullptr, 1000, &g,
                            &err) && err.find("degenerate") != std::string::npos);
    std::printf("error paths ok\n");
}

int main()
{
    check_mesh(beam(1, 1, 1, 0.0, 1), false, "one cell");
    check_mesh(beam(16, 4, 4, 0.0, 1), false, "16x4x4 beam");
    check_mesh(beam(24, 6, 8, 0.2, 7), true, "24x6x8 jittered, per-element moduli");
    // a single tetrahedron and nodes that no element references
    Mesh t;
    t.X = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 5, 5, 5, 6, 6, 6};
    t.tets = {0, 1, 2, 3};
    t.fixed = {1, 0, 0, 0, 0, 0};
    check_mesh(t, false, "single tet with two unreferenced nodes");
    // elements in random order over a shuffled node numbering
    Mesh r = beam(12, 5, 3, 0.1, 3);
    std::mt19937 rng(5);
    std::vector<int> perm(r.nn());
    for (int i = 0; i < (int)perm.size(); i++) perm[i] = i;
    std::shuffle(perm.begin(), perm.end(), rng);
    Mesh s;
    s.X.resize(r.X.size());
    s.fixed.resize(r.fixed.size());
    for (int64_t i = 0; i < r.nn(); i++) {
        for (int k = 0; k < 3; k++) s.X[3 * perm[i] + k] = r.X[3 * i + k];
        s.fixed[perm[i]] = r.fixed[i];
    }
    std::vector<int> eo(r.ne());
    for (int i = 0; i < (int)eo.size(); i++) eo[i] = i;
    std::shuffle(eo.begin(), eo.end(), rng);
    for (int e : eo)
        for (int k = 0; k < 4; k++) s.tets.push_back(perm[r.tets[4 * e + k]]);
    check_mesh(s, true, "shuffled nodes and elements");
    check_csr_helpers();
    check_multigrid_setup();
    check_errors();
    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("all host checks passed\n");
    return 0;
}
