// Host BVH build for the MI355X traversal kernels.
//
// Replaces the reference's BvhAccel build (bvhAccel.cpp:13-149: midpoint
// split, one primitive per leaf, 64-byte FP64 LinearBvhNode) with a binned-SAH
// binary build collapsed to a 4-wide tree and flattened to 128-byte SoA
// nodes, then quantized to 64-byte nodes (8-bit child planes on a per-node
// power-of-two grid), half an HBM/L2 cache line per node fetch; the quantized
// boxes contain the exact ones, so traversal only ever tests more
// triangles, never fewer.  Closest-hit results do not
// depend on tree shape: ids are mapped back to the caller's original
// triangle order, so the reference's in-place reorder (bvhAccel.cpp:114-121)
// never reaches the caller.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#if defined(__HIPCC__)
#define CHK_HD __host__ __device__
#else
#define CHK_HD
#endif

namespace chk {

// 4-wide node: child boxes stored per axis (SoA), 128 bytes, 128-aligned
// (built on the host and quantized to Node4Q for the device).
struct alignas(128) Node4 {
  float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
  uint32_t child[4];  // kLeafBit clear: node index; set: leaf (see leaf_*)
  uint32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "Node4 must be one 128-byte line");

// The same node with each child box quantized outward to an 8-bit grid per
// axis (Ylitie, Karras and Laine, "Efficient Incoherent Ray Traversal on GPUs
// Through Compressed Wide BVHs", HPG 2017): axis a's planes are
// o_a + q * s_a with s_a a power of two and q in 0..255, q rounded down for
// low planes and up for high ones, so each quantized box contains its
// child's box.  64 bytes: half the bytes per node visit.
struct alignas(64) Node4Q {
  float ox, oy, oz, sx;          // grid origin; x step
  uint32_t qlx, qhx, qly, qhy;   // byte i: child i's low / high grid index
  uint32_t qlz, qhz;
  float sy, sz;                  // y, z steps
  uint32_t child[4];             // as Node4::child
};
static_assert(sizeof(Node4Q) == 64, "Node4Q must be 64 bytes");

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kMaxLeaf = 4;
constexpr uint32_t kEmptyChild = 0xFFFFFFFFu;
CHK_HD inline uint32_t leaf_make(uint32_t first, uint32_t count) {
  return kLeafBit | ((count - 1) << 28) | first;
}
CHK_HD inline uint32_t leaf_first(uint32_t c) { return c & 0x0FFFFFFFu; }
CHK_HD inline uint32_t leaf_count(uint32_t c) { return ((c >> 28) & 7u) + 1; }

// Triangle in leaf order: three float4 (A, id), (B, -), (C, -); w of the first
// holds the ORIGINAL triangle index as raw bits.
struct alignas(16) TriRec {
  float a[4], b[4], c[4];
};

struct BvhStats {
  uint32_t n_nodes = 0;
  uint32_t n_leaves = 0;
  uint32_t n_tris = 0;        // triangles in the tree (degenerate ones dropped)
  uint32_t max_depth = 0;     // node levels
  uint32_t stack_bound = 0;   // max over root-leaf paths of sum(children-1): traversal stack need
  double sah_cost = 0;
};

struct Bvh4 {
  std::vector<Node4> nodes;
  std::vector<Node4Q> qnodes;     // nodes, quantized (what the device traverses)
  std::vector<TriRec> tris;       // leaf order
  std::vector<uint32_t> leaf_to_orig;
  std::vector<int32_t> orig_to_leaf;  // -1 for dropped (degenerate) triangles
  BvhStats stats;
};

// tris: n x 9 floats (A, B, C).  Triangles with zero or non-finite area are
// dropped: the reference's test never reports them (NaN normal, see
// geometryPrimitives.h:75-79 and prepareScene.cpp:241).
// Returns false with msg on failure.
bool build_bvh4(const float *tris, uint32_t n, Bvh4 &out, std::string *msg);

}  // namespace chk
