// sfem_host.hpp -- the host-only set-up logic of libsfem_hip.so: element
// geometry, the element-block partition, and the CSR tables (pressure
// triangles per node, restriction of a multigrid level, cable attachments).
// No HIP types: sfem_host.cpp also builds for the CPU alone, which is how it is
// run under AddressSanitizer and UndefinedBehaviorSanitizer (tests/host_check.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace sfem {

constexpr uint32_t F_SHARED = 0x80000000u;
constexpr uint32_t F_FIXED = 0x40000000u;
constexpr uint32_t F_OWNER = 0x20000000u;   // the node's only block, or the first of its blocks
constexpr uint32_t F_NODE_MASK = 0x1fffffffu;

// per element Dm^-1 (9, row-major), signed rest volume det(Dm) / 6 and Lame
// parameters (FemPhysicsMatrixFree.cpp:126-146, FemPhysicsBase.cpp:40-41);
// per node lumped mass rho V / 4 per corner, 0 on held nodes; `held` = the
// caller's fixed flags plus every node that no element references
struct ElementGeometry {
    std::vector<double> dminv, vol, mu, lam, mass;
    std::vector<uint8_t> held;
    int64_t nheld = 0;
};

// false (and *err set) for a degenerate element or an index out of range
bool element_geometry(const int32_t* tets, int64_t ne, const double* X0, int64_t nn, const uint8_t* fixed, double young,
                      double poisson, const double* young_e, const double* poisson_e, double density, ElementGeometry* out,
                      std::string* err);

// The element-block partition of the element kernels: elements sorted by the
// Morton code of their centroid and cut into blocks of block_e; per block its
// node rows (sorted node ids with flags), per element 4 block-local ids and
// per corner its slot in the block's node-sorted incidence list; nodes of
// several blocks get one partial-row slot per block, contiguous per node in
// block order.
struct Partition {
    int nblk = 0, nmax = 0;
    int64_t ne_pad = 0;
    std::vector<int> bnode_ptr;        // nblk + 1
    std::vector<uint32_t> bnodes;      // node id | F_SHARED | F_FIXED | F_OWNER per block row
    std::vector<uint16_t> lptr;        // per block nb + 1 incidence offsets
    std::vector<uint32_t> lidx, lperm; // 2 x ne_pad: 4 x 16-bit local ids / incidence slots
    std::vector<float> dm32, cf32;     // 9 x ne_pad, 2 x ne_pad (V mu, V lambda), block order
    std::vector<double> dm64, cf64;
    std::vector<float> rel32;          // 3 per block row: rest position minus the block's box centre
    bool uniform = true;               // mu, lambda the same for every element
    std::vector<float> mat32;          // 2 x ne_pad mu, lambda when not uniform
    std::vector<int> eorig;            // sorted element -> original index
    std::vector<uint32_t> sh_node;     // shared nodes: id | F_FIXED
    std::vector<int> sh_ptr, bslot;    // shared node -> slot range; block row -> slot (-1 unshared)
    std::vector<int> node_q;           // 2 nn + 2: per node {first slot, end slot}, {-1,0}, {-2,0}; then {0,0}
};

Partition make_partition(const int32_t* tets, int64_t ne, const double* X0, int64_t nn, const uint8_t* held,
                         const std::vector<double>& mu, const std::vector<double>& lam, const std::vector<double>& dminv,
                         const std::vector<double>& vol, int block_e);

// node -> incident triangles (CSR, triangle order) of an n-node mesh
void node_triangle_csr(const int32_t* tris, int64_t ntri, int64_t nn, std::vector<int>* ptr, std::vector<int>* idx);

// the restriction of a prolongation P (CSR over na rows, columns < nc): per
// coarse node its (fine row, weight) entries in increasing row order
void transpose_prolongation(const int64_t* p_ptr, const int32_t* p_col, const double* p_val, int64_t na, int64_t nc,
                            std::vector<int>* tptr, std::vector<int>* trow, std::vector<float>* tval);

// the transfer arguments of a multigrid level below a level of na nodes:
// nc coarse nodes (0 < nc < na), inject[nc] in [0, na), the prolongation as
// CSR over the na rows (p_ptr[0] = 0, monotone, nnz < 2^31, columns in
// [0, nc)); false and *err set otherwise
bool check_level_transfer(int64_t na, int64_t nc, const int32_t* inject, const int64_t* p_ptr, const int32_t* p_col,
                          std::string* err);

// the element data of a directly solved (dense) level, element order: its
// corners, Dm^-1 (FP32, row-major) and (V mu, V lambda) (FP32) as the context
// computes them at set-up (FemPhysicsMatrixFree.cpp:126-146,
// FemPhysicsBase.cpp:40-41), and per node its (element, corner) incidences
// in element order (CSR: inc_ptr[nn + 1], inc = 2 ints per incidence)
struct DenseLevelData {
    std::vector<int> tets4, inc_ptr, inc;
    std::vector<float> dminv, cvol;
};
DenseLevelData dense_level_data(const int32_t* tets, int64_t ne, const double* X0, int64_t nn, double young,
                                double poisson, const double* young_e, const double* poisson_e);

// per mesh node (of nn) the cable attachments (cable node j, weight index
// 4 j + k) of cable nodes whose element corners are node4[4 j + k]; the
// touched nodes in increasing order with CSR offsets
void cable_attachments(const std::vector<int>& node4, int64_t nn, std::vector<int>* touched, std::vector<int>* tptr,
                       std::vector<std::pair<int, int>>* tent);

}  // namespace sfem
