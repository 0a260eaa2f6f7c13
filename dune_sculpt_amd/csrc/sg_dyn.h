// Dynamic topology (PBVH_BMESH): the mesh container, the host tree, the device state.
//
// Corner table: triangle t has corners 3t..3t+2; tv[c] is the corner's vertex and
// topp[c] the corner across the edge opposite c (tv[next c] -> tv[prev c]), -1 on a
// boundary.  Every vertex has a 64-bit label (input index; a split's new vertex gets a
// hash of its edge's labels, the topology update's serial and the round), so results
// never depend on slot numbers and slots may be allocated in any order.
#pragma once

#include <cstdint>
#include <vector>

#include "sg_internal.h"

#define SG_DYN_MAX_ROUNDS 64  // rounds per phase (collapse, subdivide) of one topology update
#define SG_DYN_RING 128       // faces around a vertex that a walk follows (more: the vertex is skipped)
#define SG_DYN_STACK 32       // pending edges of the even-subdivision walk
#define SG_DYN_ANC 64         // ancestors per node in the ancestor table (a wave's lanes)

// The caller's mesh (stand-in for BMesh).  Host arrays until a PBVH is built over it;
// then the PBVH's HBM copy is the mesh (owner set) and reads go to the device.
struct SgDynMesh {
  int V = 0, T = 0, Vcap = 0, Tcap = 0;
  bool has_mask = false;
  std::vector<float> co, no, mask, fno;  // co/no [3*Vcap], fno [3*Tcap]
  std::vector<uint8_t> vhide, fhide;
  std::vector<uint64_t> lab;
  std::vector<int> vnode, vcorner, tv, topp, tnode, tslot;
  PBVH *owner = nullptr;
};

enum {
  SGD_V = 0,        // vertex slots in use
  SGD_T = 1,        // face slots in use
  SGD_NCAND = 2,    // candidates registered this round
  SGD_NVALID = 3,   // candidates that may run
  SGD_NWIN = 4,     // winners
  SGD_WN = 5,       // faces in the working set
  SGD_HV = 6,       // opposite vertices of the last round's splits (read this round)
  SGD_HV2 = 7,      // ... of this round's splits
  SGD_FLAGS = 8,    // 1: capacity stop, 2: walk stack overflow
  SGD_NGATHER = 9,  // gathered leaves
  SGD_NVL = 10,     // vertex list (brush candidates, normals)
  SGD_NAFF = 11,    // affected vertices
  SGD_AREA_X = 12,
  SGD_AREA_Y = 13,
  SGD_AREA_Z = 14,
  SGD_NAREA = 15,
  SGD_NTOUCH = 16,  // vertices a collapse round touched
  SGD_NLIST = 17,   // flagged leaves
  SGD_NFN = 18,     // faces re-normalled
  SGD_NREFIT = 19,  // leaves refit
  SGD_LN = 20,      // faces of the working set in range (the round's scan list)
  SGD_NS = 21,      // seed faces this round (in range, with an edge past the phase's limit)
  SGD_NS2 = 22,     // ... for the next round
  SGD_NLIST2 = 23,  // leaves flagged for the bounds (the normals' list is SGD_NLIST)
  SGD_ST = 24,      // topology counters: splits, collapses, split rounds, collapse rounds,
                    // split candidates, collapse candidates, leaves in the working set
  SGD_SERIAL = 31,  // topology updates so far (labels of new vertices hash it)
  SGD_COUNT = 32,
};

struct SgDynDev {
  int Vcap = 0, Tcap = 0, Ncap = 0, N = 0, L = 0, H = 0;
  int has_mask = 0, ccap = 0;
  // vertices
  float *px = nullptr, *py = nullptr, *pz = nullptr, *nx = nullptr, *ny = nullptr, *nz = nullptr, *mask = nullptr;
  uint8_t *vhide = nullptr;
  unsigned long long *lab = nullptr;
  int *vnode = nullptr, *vcorner = nullptr, *vstamp = nullptr, *astamp = nullptr, *aslot = nullptr, *hvst = nullptr;
  // faces
  int *tv = nullptr, *topp = nullptr, *tnode = nullptr, *tslot = nullptr, *wstamp = nullptr;
  unsigned long long *tie = nullptr;  // largest key two edges share on the face
  float *fnx = nullptr, *fny = nullptr, *fnz = nullptr;
  uint8_t *fhide = nullptr;
  unsigned long long *claim = nullptr;
  // corners
  int *estamp = nullptr, *remap = nullptr, *remap_ser = nullptr;
  // nodes
  int *node_flag = nullptr, *node_child = nullptr, *node_dirty = nullptr, *leaf_node = nullptr;
  int *lvl_off = nullptr, *lvl_node = nullptr;  // internal nodes grouped by height (children first)
  int *node_parent = nullptr, *node_height = nullptr, *anc_stamp = nullptr, *anc_pos = nullptr;
  int *anc = nullptr;  // [N * SG_DYN_ANC]: a node's ancestors, parent first (-1 past the root)
  int *lvl_list = nullptr, *lvl_cnt = nullptr;  // ancestors of refit leaves, by height (lvl_off layout)
  float *node_bb = nullptr, *node_obb = nullptr;  // [8] per node: min xyz, pad, max xyz, pad
  int *slab = nullptr, *scount = nullptr;
  // work lists
  int *wlist = nullptr, *cand = nullptr, *win = nullptr, *outer = nullptr, *hv0 = nullptr, *hv1 = nullptr;
  int *llist = nullptr, *lstamp = nullptr;  // in-range faces of the working set, membership stamp
  int *seed0 = nullptr, *seed1 = nullptr, *sstamp = nullptr;  // seed faces of a round, the next round's, stamp
  uint8_t *fin = nullptr;                    // per face: in range (kept exact through the rounds)
  int *touched = nullptr, *gathered = nullptr, *vlist = nullptr, *aff = nullptr, *nlist = nullptr, *nlist2 = nullptr;
  unsigned long long *key = nullptr;
  uint8_t *valid = nullptr;
  float *fade = nullptr, *buf0 = nullptr, *buf1 = nullptr;
  float *curve = nullptr;
  long long *ctl = nullptr;
  long long *ctl_host = nullptr;  // pinned
};

// Host tree of a PBVH_BMESH (authoritative at build and after_stroke, else a mirror).
struct SgDynTree {
  std::vector<int> child, flag, scount, slab;  // slab [N * SG_DYN_SLAB]
  std::vector<float> bb, obb;                  // [N * 8]
};

struct SgDynState {
  SgDynMesh *bm = nullptr;
  SgDynDev d;
  SgDynTree t;
  float max_len = 0.75f, min_len = 0.3f;
  int mode = PBVH_Subdivide | PBVH_Collapse, use_frontface = 0, use_projected = 0;
  float view_normal[3] = {0.0f, 0.0f, 1.0f};
  int serial = 0;        // topology updates so far (HBM holds the live count while uploaded)
  int stamp = 0;         // round / dedupe stamps
  int tag_base = 0;      // claims carry (round stamp - tag_base) above their key
  long long stats[12] = {0};
  bool host_stale = false;  // the host tree mirror trails the device
  std::vector<int> pend;    // node flag bits set by host calls, not yet on the device
};

// host (dyntopo_api.cpp)
int sg_dynmesh_build_tables(SgDynMesh *m);  // corner table, owners reset, normals
void sg_dyn_build_tree(const SgDynMesh &m, int leaf_limit, SgDynTree &t, std::vector<int> &vnode,
                       std::vector<int> &tnode, std::vector<int> &tslot);
void sg_dyn_after_stroke_host(SgDynMesh &m, int leaf_limit, SgDynTree &t);

int sg_dyn_refresh(PBVH *p);                    // host tree mirror from HBM when stale
void sg_dyn_note_flags(PBVHNode *node, int bits);  // host-set node flags for the device
void sg_dyn_release(PBVH *p);                    // device state and the mesh link (BKE_pbvh_free)
int sg_dyn_last_gathered(PBVH *p, int *out, int cap);

// device (dyntopo.hip)
int sg_dyn_upload(PBVH *p);                 // mesh + tree -> HBM (allocates on first use)
int sg_dyn_download(PBVH *p, bool mesh);    // HBM -> host mesh (mesh) and tree
void sg_dyn_free_device(PBVH *p);
int sg_dyn_update_topology(PBVH *p, int mode, const float center[3], float radius);  // waits; 1 if modified
int sg_dyn_dab_apply(PBVH *p, const SgDabParams *dab);
int sg_dyn_update_normals_bounds(PBVH *p, bool normals, bool bounds);
int sg_dyn_sync_flags_down(PBVH *p);
int sg_dyn_last_affected(PBVH *p, uint64_t *out, int cap);
