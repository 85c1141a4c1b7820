// sfem_host.cpp -- host-only set-up logic of libsfem_hip.so (see sfem_host.hpp).
#include "sfem_host.hpp"

#include <algorithm>
#include <cmath>

namespace sfem {

namespace {

// 21-bit-per-axis Morton code
uint64_t spread21(uint64_t v)
{
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

}  // namespace

bool element_geometry(const int32_t* tets, int64_t ne, const double* X0, int64_t nn, const uint8_t* fixed, double young,
                      double poisson, const double* young_e, const double* poisson_e, double density, ElementGeometry* g,
                      std::string* err)
{
    g->dminv.assign(9 * (size_t)ne, 0.0);
    g->vol.assign(ne, 0.0);
    g->mu.assign(ne, 0.0);
    g->lam.assign(ne, 0.0);
    std::vector<double> msum(nn, 0.0);
    for (int64_t e = 0; e < ne; e++) {
        const int32_t* t = tets + 4 * e;
        for (int k = 0; k < 4; k++)
            if (t[k] < 0 || t[k] >= nn) {
                *err = "node index out of range";
                return false;
            }
        double D[9];
        for (int k = 1; k < 4; k++)
            for (int r = 0; r < 3; r++) D[3 * r + (k - 1)] = X0[3 * (int64_t)t[k] + r] - X0[3 * (int64_t)t[0] + r];
        // adjugate inverse (Math/matrix3.h:234-249)
        double v[9];
        v[0] = D[4] * D[8] - D[5] * D[7];
        v[3] = -D[3] * D[8] + D[6] * D[5];
        v[6] = D[3] * D[7] - D[6] * D[4];
        v[1] = D[2] * D[7] - D[1] * D[8];
        v[4] = D[0] * D[8] - D[2] * D[6];
        v[7] = D[1] * D[6] - D[0] * D[7];
        v[2] = D[1] * D[5] - D[2] * D[4];
        v[5] = D[2] * D[3] - D[0] * D[5];
        v[8] = D[0] * D[4] - D[1] * D[3];
        const double det = D[0] * v[0] + D[1] * v[3] + D[2] * v[6];
        if (det == 0.0) {
            *err = "degenerate element " + std::to_string(e);
            return false;
        }
        const double s = 1.0 / det;
        for (int k = 0; k < 9; k++) g->dminv[9 * e + k] = v[k] * s;
        g->vol[e] = det / 6.0;
        const double E = young_e ? young_e[e] : young;
        const double nu = poisson_e ? poisson_e[e] : poisson;
        g->lam[e] = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu);
        g->mu[e] = 0.5 * E / (1.0 + nu);
        const double lm = 0.25 * g->vol[e] * density;
        for (int k = 0; k < 4; k++) msum[t[k]] += lm;
    }
    // a node that no element references has neither stiffness nor mass: it is
    // held like a fixed node (its block would be singular)
    g->held.assign(fixed, fixed + nn);
    std::vector<uint8_t> used(nn, 0);
    for (int64_t k = 0; k < 4 * ne; k++) used[tets[k]] = 1;
    for (int64_t i = 0; i < nn; i++)
        if (!used[i]) g->held[i] = 1;
    g->mass.assign(nn, 0.0);
    g->nheld = 0;
    for (int64_t i = 0; i < nn; i++) {
        if (g->held[i]) {
            g->nheld++;
            continue;
        }
        const double inv = msum[i] != 0 ? 1.0 / msum[i] : 1.0;   // Node::invMass default 1
        g->mass[i] = 1.0 / inv;
    }
    return true;
}

// LDS bank-aware order of a block's nodes.  The Hessian-vector kernel stages
// one 16-byte row per block node in LDS and each lane (one element) gathers
// the rows of its four corners; rows whose index agrees modulo ROW_CLASSES
// fall in the same LDS banks, and distinct such rows read by the lanes of one
// group of GROUP_LANES in one instruction are served one after the other.
// The sorted order puts them there at random.  This order colours the block's
// nodes with the ROW_CLASSES residues (each residue as often as the block has
// rows for it) so that the distinct nodes a lane group reads at one corner get
// distinct residues wherever the greedy colouring finds one (most constrained
// node first); a node's residue class is then filled in sorted order.
// -> order[new local id] = index into the sorted ids.  Large blocks (more than
// MAX_ORDERED nodes, unstructured meshes) keep the sorted order.
#ifndef SFEM_ROW_CLASSES
#define SFEM_ROW_CLASSES 8
#endif
#ifndef SFEM_GROUP_LANES
#define SFEM_GROUP_LANES 8
#endif
static std::vector<int> bank_order(const std::vector<int>& lsorted, int64_t nel, int nb)
{
    constexpr int R = SFEM_ROW_CLASSES, L = SFEM_GROUP_LANES, MAX_ORDERED = 1024;
    std::vector<int> order(nb);
    for (int i = 0; i < nb; i++) order[i] = i;
    if (nb > MAX_ORDERED || nb <= R) return order;
    // w[a * nb + b]: how often nodes a and b are read by the same corner of
    // one lane group
    std::vector<uint16_t> w((size_t)nb * nb, 0);
    std::vector<int> set;
    for (int64_t g0 = 0; g0 < nel; g0 += L)
        for (int v = 0; v < 4; v++) {
            set.clear();
            for (int64_t s = g0; s < std::min<int64_t>(nel, g0 + L); s++) set.push_back(lsorted[4 * s + v]);
            std::sort(set.begin(), set.end());
            set.erase(std::unique(set.begin(), set.end()), set.end());
            for (size_t i = 0; i < set.size(); i++)
                for (size_t j = i + 1; j < set.size(); j++) {
                    w[(size_t)set[i] * nb + set[j]]++;
                    w[(size_t)set[j] * nb + set[i]]++;
                }
        }
    std::vector<int> colour(nb, -1), cap(R), deg(nb, 0);
    for (int c = 0; c < R; c++) cap[c] = (nb - c + R - 1) / R;
    for (int a = 0; a < nb; a++)
        for (int b = 0; b < nb; b++) deg[a] += w[(size_t)a * nb + b];
    // cost[a * R + c]: weight of a's coloured neighbours of colour c
    std::vector<int> cost((size_t)nb * R, 0), sat(nb, 0);
    for (int step = 0; step < nb; step++) {
        // the uncoloured node with the most distinct neighbour colours, then the largest weight
        int a = -1;
        for (int i = 0; i < nb; i++)
            if (colour[i] < 0 && (a < 0 || sat[i] > sat[a] || (sat[i] == sat[a] && deg[i] > deg[a]))) a = i;
        int best = -1;
        for (int c = 0; c < R; c++)
            if (cap[c] > 0 && (best < 0 || cost[(size_t)a * R + c] < cost[(size_t)a * R + best] ||
                               (cost[(size_t)a * R + c] == cost[(size_t)a * R + best] && cap[c] > cap[best])))
                best = c;
        colour[a] = best;
        cap[best]--;
        for (int b = 0; b < nb; b++) {
            const int wab = w[(size_t)a * nb + b];
            if (!wab || colour[b] >= 0) continue;
            if (cost[(size_t)b * R + best] == 0) sat[b]++;
            cost[(size_t)b * R + best] += wab;
        }
    }
    // residue class c holds local ids c, c + R, c + 2R, ...
    std::vector<int> next(R);
    for (int c = 0; c < R; c++) next[c] = c;
    for (int i = 0; i < nb; i++) {
        order[next[colour[i]]] = i;
        next[colour[i]] += R;
    }
    return order;
}

// extra cycles of one lane group's read of rows l[0..n) (distinct rows of one
// residue class are served one after the other)
static int group_conflicts(const int* l, int n)
{
    constexpr int R = SFEM_ROW_CLASSES;
    int cnt[R] = {0}, best = 0;
    for (int i = 0; i < n; i++) {
        bool dup = false;
        for (int k = 0; k < i; k++) dup = dup || l[k] == l[i];
        if (!dup) best = std::max(best, ++cnt[l[i] % R]);
    }
    return best > 0 ? best - 1 : 0;
}

// The elements of one block in lane groups: each group of GROUP_LANES
// consecutive elements reads its four corners' staged rows one corner per
// instruction, so a group's rows at one corner should fall in distinct
// residue classes.  Greedy: each lane takes, among the next WINDOW unplaced
// elements, the first that adds the fewest conflicts.  lid4: the elements'
// local ids (4 per element, in their current order) -> perm[new position] =
// current position.
static std::vector<int> lane_groups(const std::vector<int>& lid4, int n)
{
    constexpr int L = SFEM_GROUP_LANES, WINDOW = 64;
    std::vector<int> perm, rest(n);
    perm.reserve(n);
    for (int i = 0; i < n; i++) rest[i] = i;
    while (!rest.empty()) {
        int grp[L], m = 0;
        while (m < L && !rest.empty()) {
            int best = 0, bc = 1 << 30;
            for (int c = 0; c < (int)rest.size() && c < WINDOW; c++) {
                int add = 0;
                for (int v = 0; v < 4; v++) {
                    int l[L];
                    for (int i = 0; i < m; i++) l[i] = lid4[4 * grp[i] + v];
                    const int before = group_conflicts(l, m);
                    l[m] = lid4[4 * rest[c] + v];
                    add += group_conflicts(l, m + 1) - before;
                }
                if (add < bc) {
                    bc = add;
                    best = c;
                    if (add == 0) break;
                }
            }
            grp[m++] = rest[best];
            perm.push_back(rest[best]);
            rest.erase(rest.begin() + best);
        }
    }
    return perm;
}

Partition make_partition(const int32_t* tets, int64_t ne, const double* X0, int64_t nn, const uint8_t* held,
                         const std::vector<double>& mu, const std::vector<double>& lam, const std::vector<double>& dminv,
                         const std::vector<double>& vol, int BE)
{
    Partition P;
    // --- space-filling-curve order of elements (centroid Morton code, cubic cells)
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < nn; i++)
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], X0[3 * i + k]);
            hi[k] = std::max(hi[k], X0[3 * i + k]);
        }
    double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    if (!(ext > 0)) ext = 1;
    const double scale = (double)((1 << 21) - 1) / ext;
    std::vector<std::pair<uint64_t, int64_t>> key(ne);
    for (int64_t e = 0; e < ne; e++) {
        double cc[3] = {0, 0, 0};
        for (int v = 0; v < 4; v++)
            for (int k = 0; k < 3; k++) cc[k] += 0.25 * X0[3 * (int64_t)tets[4 * e + v] + k];
        uint64_t q[3];
        for (int k = 0; k < 3; k++) q[k] = (uint64_t)std::max(0.0, std::min((cc[k] - lo[k]) * scale, (double)((1 << 21) - 1)));
        key[e] = {spread21(q[0]) | spread21(q[1]) << 1 | spread21(q[2]) << 2, e};
    }
    std::sort(key.begin(), key.end());

    const int nblk = (int)((ne + BE - 1) / BE);
    const int64_t ne_pad = (int64_t)nblk * BE;
    P.nblk = nblk;
    P.ne_pad = ne_pad;
    P.bnode_ptr.assign(nblk + 1, 0);
    P.lidx.assign((size_t)ne_pad * 2, 0);
    P.lperm.assign((size_t)ne_pad * 2, 0);
    P.dm32.assign((size_t)9 * ne_pad, 0.f);
    P.cf32.assign((size_t)2 * ne_pad, 0.f);
    P.dm64.assign((size_t)9 * ne_pad, 0.0);
    P.cf64.assign((size_t)2 * ne_pad, 0.0);
    std::vector<int> node_blocks(nn, 0);
    P.bnodes.reserve((size_t)ne);
    P.lptr.reserve((size_t)ne);
    std::vector<int> ids, cnt, lsorted, lid;
    for (int b = 0; b < nblk; b++) {
        const int64_t e0 = (int64_t)b * BE;
        const int64_t e1 = std::min<int64_t>(ne, e0 + BE);
        ids.clear();
        for (int64_t s = e0; s < e1; s++)
            for (int v = 0; v < 4; v++) ids.push_back(tets[4 * key[s].second + v]);
        std::sort(ids.begin(), ids.end());
        ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
        const int nb = (int)ids.size();
        P.nmax = std::max(P.nmax, nb);
        P.bnode_ptr[b + 1] = P.bnode_ptr[b] + nb;
        // the block's rows in LDS bank-aware order: local id of sorted index i = lid[i]
        lsorted.assign((size_t)4 * (e1 - e0), 0);
        for (int64_t s = e0; s < e1; s++)
            for (int v = 0; v < 4; v++)
                lsorted[4 * (s - e0) + v] =
                    (int)(std::lower_bound(ids.begin(), ids.end(), tets[4 * key[s].second + v]) - ids.begin());
        const std::vector<int> order = bank_order(lsorted, e1 - e0, nb);
        lid.assign(nb, 0);
        for (int j = 0; j < nb; j++) lid[order[j]] = j;
        // the block's elements in conflict-free lane groups where possible
        {
            const int n = (int)(e1 - e0);
            std::vector<int> lid4(4 * (size_t)n);
            for (int i = 0; i < 4 * n; i++) lid4[i] = lid[lsorted[i]];
            const std::vector<int> perm = lane_groups(lid4, n);
            const std::vector<std::pair<uint64_t, int64_t>> kb(key.begin() + e0, key.begin() + e1);
            const std::vector<int> lsb(lsorted);
            for (int i = 0; i < n; i++) {
                key[e0 + i] = kb[perm[i]];
                for (int v = 0; v < 4; v++) lsorted[4 * i + v] = lsb[4 * perm[i] + v];
            }
        }
        for (int j = 0; j < nb; j++) {
            const int id = ids[order[j]];
            P.bnodes.push_back((uint32_t)id | (held[id] ? F_FIXED : 0u));
            node_blocks[id]++;
        }
        // local ids and the per-node incidence lists (counting sort, element order)
        cnt.assign(nb + 1, 0);
        for (int64_t s = e0; s < e1; s++) {
            const int64_t e = key[s].second;
            int l[4];
            for (int v = 0; v < 4; v++) {
                l[v] = lid[lsorted[4 * (s - e0) + v]];
                cnt[l[v] + 1]++;
            }
            P.lidx[2 * s] = (uint32_t)l[0] | (uint32_t)l[1] << 16;
            P.lidx[2 * s + 1] = (uint32_t)l[2] | (uint32_t)l[3] << 16;
            for (int k = 0; k < 9; k++) {
                P.dm64[(size_t)k * ne_pad + s] = dminv[9 * e + k];
                P.dm32[(size_t)k * ne_pad + s] = (float)dminv[9 * e + k];
            }
            const double cm = vol[e] * mu[e], cl = vol[e] * lam[e];
            P.cf64[s] = cm;
            P.cf64[ne_pad + s] = cl;
            P.cf32[s] = (float)cm;
            P.cf32[ne_pad + s] = (float)cl;
        }
        for (int j = 0; j < nb; j++) cnt[j + 1] += cnt[j];
        for (int j = 0; j <= nb; j++) P.lptr.push_back((uint16_t)cnt[j]);
        // corner v of element s -> a slot of its node's range in the node-sorted
        // incidence order.  The corner writes of one lane group (one corner per
        // instruction) are served together: each takes the first free slot of
        // its node whose residue class the group has not written at that corner
        // yet, else the first free slot.
        std::vector<std::vector<int>> free_slots(nb);
        for (int j = 0; j < nb; j++)
            for (int q = cnt[j]; q < cnt[j + 1]; q++) free_slots[j].push_back(q);
        for (int64_t g0 = e0; g0 < e1; g0 += SFEM_GROUP_LANES) {
            const int64_t g1 = std::min<int64_t>(e1, g0 + SFEM_GROUP_LANES);
            uint16_t k[SFEM_GROUP_LANES][4];
            for (int v = 0; v < 4; v++) {
                unsigned used = 0;
                for (int64_t s = g0; s < g1; s++) {
                    const uint32_t w = P.lidx[2 * s + v / 2];
                    const int l = (int)(v % 2 ? w >> 16 : w & 0xffff);
                    std::vector<int>& F = free_slots[l];
                    size_t pick = 0;
                    while (pick < F.size() && (used >> (F[pick] % SFEM_ROW_CLASSES) & 1u)) pick++;
                    if (pick == F.size()) pick = 0;
                    const int q = F[pick];
                    F.erase(F.begin() + (long)pick);
                    used |= 1u << (q % SFEM_ROW_CLASSES);
                    k[s - g0][v] = (uint16_t)q;
                }
            }
            for (int64_t s = g0; s < g1; s++) {
                P.lperm[2 * s] = (uint32_t)k[s - g0][0] | (uint32_t)k[s - g0][1] << 16;
                P.lperm[2 * s + 1] = (uint32_t)k[s - g0][2] | (uint32_t)k[s - g0][3] << 16;
            }
        }
    }
    // shared nodes: flag, and list their partial rows in block order
    std::vector<int> sh_of(nn, -1);
    for (int64_t i = 0; i < nn; i++)
        if (node_blocks[i] > 1) {
            sh_of[i] = (int)P.sh_node.size();
            P.sh_node.push_back((uint32_t)i | (held[i] ? F_FIXED : 0u));
        }
    P.sh_ptr.assign(P.sh_node.size() + 1, 0);
    for (size_t r = 0; r < P.bnodes.size(); r++) {
        const uint32_t id = P.bnodes[r] & F_NODE_MASK;
        if (sh_of[id] >= 0) {
            P.bnodes[r] |= F_SHARED;
            P.sh_ptr[sh_of[id] + 1]++;
        }
    }
    for (size_t i = 0; i < P.sh_node.size(); i++) P.sh_ptr[i + 1] += P.sh_ptr[i];
    // each shared node's partial rows get contiguous slots, in block order
    P.bslot.assign(P.bnodes.size(), -1);
    std::vector<int> fill(P.sh_ptr.begin(), P.sh_ptr.end() - 1);
    for (size_t r = 0; r < P.bnodes.size(); r++) {   // rows are in block order
        const uint32_t id = P.bnodes[r] & F_NODE_MASK;
        if (sh_of[id] >= 0) P.bslot[r] = fill[sh_of[id]]++;
        // the row that stores a node's per-node values: its only block's, or
        // its first block's (the one with the node's first partial slot)
        if (sh_of[id] < 0 || P.bslot[r] == P.sh_ptr[sh_of[id]]) P.bnodes[r] |= F_OWNER;
    }
    // rest coordinates of every block row relative to the centre of the block's
    // bounding box (FP64 difference, then FP32): |coordinate| <= half the block
    // diameter, so the FP32 edge vectors the Hessian-vector kernel forms from
    // them are accurate to ~2^-24 block diameter / edge
    P.rel32.assign(3 * P.bnodes.size(), 0.f);
    for (int b = 0; b < nblk; b++) {
        double blo[3] = {1e300, 1e300, 1e300}, bhi[3] = {-1e300, -1e300, -1e300};
        for (int r = P.bnode_ptr[b]; r < P.bnode_ptr[b + 1]; r++) {
            const int64_t id = P.bnodes[r] & F_NODE_MASK;
            for (int k = 0; k < 3; k++) {
                blo[k] = std::min(blo[k], X0[3 * id + k]);
                bhi[k] = std::max(bhi[k], X0[3 * id + k]);
            }
        }
        for (int r = P.bnode_ptr[b]; r < P.bnode_ptr[b + 1]; r++) {
            const int64_t id = P.bnodes[r] & F_NODE_MASK;
            for (int k = 0; k < 3; k++) P.rel32[3 * (size_t)r + k] = (float)(X0[3 * id + k] - 0.5 * (blo[k] + bhi[k]));
        }
    }
    // Lame parameters for the Hessian-vector kernel: per element only when they vary
    P.uniform = true;
    for (int64_t e = 1; e < ne && P.uniform; e++) P.uniform = mu[e] == mu[0] && lam[e] == lam[0];
    if (!P.uniform) {
        P.mat32.assign((size_t)2 * ne_pad, 0.f);
        for (int64_t s2 = 0; s2 < ne; s2++) {
            P.mat32[s2] = (float)mu[key[s2].second];
            P.mat32[ne_pad + s2] = (float)lam[key[s2].second];
        }
    }
    P.eorig.resize(ne);
    for (int64_t s2 = 0; s2 < ne; s2++) P.eorig[s2] = (int)key[s2].second;
    // per node, for the CG update: a shared node's partial slot range {first,
    // end}; {-1, 0} for a free node of one block; {-2, 0} for a held node and
    // for a node that no element references (its row of K is 0); after the nn
    // pairs, one more pair that stays 0 (a flag that is never set)
    P.node_q.assign(2 * nn + 2, 0);
    for (int64_t i = 0; i < nn; i++) {
        if (held[i] || node_blocks[i] == 0) P.node_q[2 * i] = -2;
        else if (sh_of[i] < 0) P.node_q[2 * i] = -1;
        else {
            P.node_q[2 * i] = P.sh_ptr[sh_of[i]];
            P.node_q[2 * i + 1] = P.sh_ptr[sh_of[i] + 1];
        }
    }
    return P;
}

void node_triangle_csr(const int32_t* tris, int64_t ntri, int64_t nn, std::vector<int>* ptr, std::vector<int>* idx)
{
    ptr->assign(nn + 1, 0);
    for (int64_t i = 0; i < 3 * ntri; i++) (*ptr)[tris[i] + 1]++;
    for (int64_t i = 0; i < nn; i++) (*ptr)[i + 1] += (*ptr)[i];
    idx->assign(ptr->back(), 0);
    std::vector<int> fill(ptr->begin(), ptr->end() - 1);
    for (int64_t k = 0; k < ntri; k++)
        for (int v = 0; v < 3; v++) (*idx)[fill[tris[3 * k + v]]++] = (int)k;
}

void transpose_prolongation(const int64_t* p_ptr, const int32_t* p_col, const double* p_val, int64_t na, int64_t nc,
                            std::vector<int>* tptr, std::vector<int>* trow, std::vector<float>* tval)
{
    const int64_t nnz = p_ptr[na];
    tptr->assign(nc + 1, 0);
    trow->assign(nnz, 0);
    tval->assign(nnz, 0.f);
    for (int64_t k = 0; k < nnz; k++) (*tptr)[p_col[k] + 1]++;
    for (int64_t j = 0; j < nc; j++) (*tptr)[j + 1] += (*tptr)[j];
    std::vector<int> fill(tptr->begin(), tptr->end() - 1);
    for (int64_t i = 0; i < na; i++)
        for (int64_t k = p_ptr[i]; k < p_ptr[i + 1]; k++) {
            const int slot = fill[p_col[k]]++;
            (*trow)[slot] = (int)i;
            (*tval)[slot] = (float)p_val[k];
        }
}

bool check_level_transfer(int64_t na, int64_t nc, const int32_t* inject, const int64_t* p_ptr, const int32_t* p_col,
                          std::string* err)
{
    if (nc <= 0 || nc >= na) {
        *err = "a coarse level has fewer nodes than the one above";
        return false;
    }
    for (int64_t k = 0; k < nc; k++)
        if (inject[k] < 0 || inject[k] >= na) {
            *err = "injection index out of range";
            return false;
        }
    if (p_ptr[0] != 0) {
        *err = "p_ptr[0] != 0";
        return false;
    }
    for (int64_t i = 0; i < na; i++)
        if (p_ptr[i + 1] < p_ptr[i]) {
            *err = "p_ptr not monotone";
            return false;
        }
    const int64_t nnz = p_ptr[na];
    if (nnz >= (int64_t)1 << 31) {
        *err = "prolongation too large";
        return false;
    }
    for (int64_t k = 0; k < nnz; k++)
        if (p_col[k] < 0 || p_col[k] >= nc) {
            *err = "prolongation column out of range";
            return false;
        }
    return true;
}

DenseLevelData dense_level_data(const int32_t* tets, int64_t ne, const double* X0, int64_t nn, double young,
                                double poisson, const double* young_e, const double* poisson_e)
{
    DenseLevelData o;
    o.tets4.assign(4 * ne, 0);
    o.dminv.assign(9 * ne, 0.f);
    o.cvol.assign(2 * ne, 0.f);
    std::vector<int> count(nn + 1, 0);
    for (int64_t e = 0; e < ne; e++) {
        const int32_t* t = tets + 4 * e;
        double D[9], v[9];
        for (int k = 1; k < 4; k++)
            for (int r = 0; r < 3; r++) D[3 * r + (k - 1)] = X0[3 * (int64_t)t[k] + r] - X0[3 * (int64_t)t[0] + r];
        v[0] = D[4] * D[8] - D[5] * D[7];
        v[3] = -D[3] * D[8] + D[6] * D[5];
        v[6] = D[3] * D[7] - D[6] * D[4];
        v[1] = D[2] * D[7] - D[1] * D[8];
        v[4] = D[0] * D[8] - D[2] * D[6];
        v[7] = D[1] * D[6] - D[0] * D[7];
        v[2] = D[1] * D[5] - D[2] * D[4];
        v[5] = D[2] * D[3] - D[0] * D[5];
        v[8] = D[0] * D[4] - D[1] * D[3];
        const double det = D[0] * v[0] + D[1] * v[3] + D[2] * v[6];
        const double E = young_e ? young_e[e] : young, nu = poisson_e ? poisson_e[e] : poisson;
        const double lam = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu), mu = 0.5 * E / (1.0 + nu);
        for (int k = 0; k < 9; k++) o.dminv[9 * e + k] = (float)(v[k] / det);
        o.cvol[2 * e] = (float)(det / 6.0 * mu);
        o.cvol[2 * e + 1] = (float)(det / 6.0 * lam);
        for (int a = 0; a < 4; a++) {
            o.tets4[4 * e + a] = t[a];
            count[t[a] + 1]++;
        }
    }
    o.inc_ptr.assign(nn + 1, 0);
    for (int64_t i = 0; i < nn; i++) o.inc_ptr[i + 1] = o.inc_ptr[i] + count[i + 1];
    o.inc.assign(2 * (size_t)o.inc_ptr[nn], 0);
    std::vector<int> fill(o.inc_ptr.begin(), o.inc_ptr.end() - 1);
    for (int64_t e = 0; e < ne; e++)
        for (int a = 0; a < 4; a++) {
            const int slot = fill[tets[4 * e + a]]++;
            o.inc[2 * (size_t)slot] = (int)e;
            o.inc[2 * (size_t)slot + 1] = a;
        }
    return o;
}

void cable_attachments(const std::vector<int>& node4, int64_t nn, std::vector<int>* touched, std::vector<int>* tptr,
                       std::vector<std::pair<int, int>>* tent)
{
    const int n = (int)(node4.size() / 4);
    std::vector<std::vector<std::pair<int, int>>> att(nn);
    for (int j = 0; j < n; j++)
        for (int k = 0; k < 4; k++) att[node4[4 * j + k]].push_back({j, 4 * j + k});
    touched->clear();
    tent->clear();
    tptr->assign(1, 0);
    for (int64_t i = 0; i < nn; i++)
        if (!att[i].empty()) {
            touched->push_back((int)i);
            tent->insert(tent->end(), att[i].begin(), att[i].end());
            tptr->push_back((int)tent->size());
        }
}

}  // namespace sfem
