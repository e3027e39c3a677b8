// xf_kernels.h -- device-side types of the gfx950 FDTD engine.
//
// Layout in HBM.  All field components share one padded node grid of
// (nx+1) x (ny+1) x pz values (3-D; z fastest, pz = nz+1 rounded up to a
// 128-byte multiple) or (nx+1) x py values (2-D TMz; y fastest), so one
// linear index addresses Ex..Hz of the same (i, j, k) and every wave64 load
// is a contiguous, aligned row segment.  Yee coefficients are compressed to
// a per-cell uint8/uint16 index into a per-pass table of 9 values (the three
// (self, curl_a, curl_b) triples of the pass), which replaces the 18 dense
// per-cell arrays of the reference (fdtd_coefficient.cuh:149-167).
#ifndef XF_KERNELS_H_
#define XF_KERNELS_H_

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include "../../include/xfdtd_hip.h"

namespace xf {

constexpr int kMaxPml = 6;
constexpr int kMaxTfsfJobs = 12;

// Every device buffer of the engine is global memory.  A pointer read from
// memory (a batch's per-grid descriptors) is not known to the compiler to
// be global, so it would be accessed with flat instructions; a flat load
// counts on the LDS/scalar counter as well, and every wait for a scalar load
// then also waits for the field loads in flight.  gp() states the address
// space, so such accesses become global loads and stores.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gp(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

// Storage <-> compute conversion.  FP64 and FP32 store what they compute;
// bf16 stores value * scale (scale per E/H) and computes in float.
template <typename S>
struct Store;
// ld_nt / st_nt are the non-temporal forms, used for the streams a pass
// reads once and rewrites (the updated component itself), so they do not
// displace the neighbour rows other waves still need from L2.
template <>
struct Store<double> {
  using C = double;
  static __device__ __forceinline__ C ld(const double* p, int64_t i, C) { return gp(p)[i]; }
  static __device__ __forceinline__ void st(double* p, int64_t i, C v, C) { gp(p)[i] = v; }
  static __device__ __forceinline__ C ld_nt(const double* p, int64_t i, C) {
    return __builtin_nontemporal_load(gp(p) + i);
  }
  static __device__ __forceinline__ void st_nt(double* p, int64_t i, C v, C) {
    __builtin_nontemporal_store(v, gp(p) + i);
  }
};
template <>
struct Store<float> {
  using C = float;
  static __device__ __forceinline__ C ld(const float* p, int64_t i, C) { return gp(p)[i]; }
  static __device__ __forceinline__ void st(float* p, int64_t i, C v, C) { gp(p)[i] = v; }
  static __device__ __forceinline__ C ld_nt(const float* p, int64_t i, C) {
    return __builtin_nontemporal_load(gp(p) + i);
  }
  static __device__ __forceinline__ void st_nt(float* p, int64_t i, C v, C) {
    __builtin_nontemporal_store(v, gp(p) + i);
  }
};
template <>
struct Store<__hip_bfloat16> {
  using C = float;
  static __device__ __forceinline__ C ld(const __hip_bfloat16* p, int64_t i, C inv) {
    return (float)gp(reinterpret_cast<const __bf16*>(p))[i] * inv;
  }
  static __device__ __forceinline__ void st(__hip_bfloat16* p, int64_t i, C v, C scale) {
    p[i] = __float2bfloat16(v * scale);  // generic store: a global-space one costs a register
  }
  // the streaming slots use plain accesses: non-temporal 16-bit loads and
  // stores measured slower at 536^3 (59.9 against 61.9 Gcells/s)
  static __device__ __forceinline__ C ld_nt(const __hip_bfloat16* p, int64_t i, C inv) {
    return ld(p, i, inv);
  }
  static __device__ __forceinline__ void st_nt(__hip_bfloat16* p, int64_t i, C v, C scale) {
    st(p, i, v, scale);
  }
};

template <>
struct Store<__half> {
  using C = float;
  static __device__ __forceinline__ C ld(const __half* p, int64_t i, C inv) {
    return (float)gp(reinterpret_cast<const _Float16*>(p))[i] * inv;
  }
  static __device__ __forceinline__ void st(__half* p, int64_t i, C v, C scale) {
    p[i] = __float2half(v * scale);  // generic store: a global-space one costs a register
  }
  // plain accesses in the streaming slots, as for bf16 (non-temporal 16-bit
  // accesses measured slower: 57.9 against 59.9 Gcells/s at 536^3)
  static __device__ __forceinline__ C ld_nt(const __half* p, int64_t i, C inv) { return ld(p, i, inv); }
  static __device__ __forceinline__ void st_nt(__half* p, int64_t i, C v, C scale) { st(p, i, v, scale); }
};

// Compensated bf16 storage: a value y is kept as hi = bf16(y) (round to
// nearest even) and the residual y - hi, which is exact in float and at most
// half a bf16 ulp, quantised to an int8 in steps of 2^-15 of hi's power of
// two.  hi + q 2^(e-15) has at most 16 significant bits and is rebuilt
// exactly in float, so a stored value carries about 16 bits of mantissa in
// 3 bytes.  Values sit in blocks of 64: 64 bf16 words (128 B), then 64
// residual bytes (64 B), so a wave's 64 consecutive values are two
// contiguous segments.  The tag type is never dereferenced.
struct Bf16c {
  unsigned char b[3];
};

__host__ __device__ __forceinline__ void bf16c_encode(float y, uint16_t& h, int8_t& l) {
  const uint32_t yb = __builtin_bit_cast(uint32_t, y);
  const uint32_t hb = (yb + 0x7fffu + ((yb >> 16) & 1u)) & 0xffff0000u;
  h = static_cast<uint16_t>(hb >> 16);
  const float hi = __builtin_bit_cast(float, hb);
  const float p2 = __builtin_bit_cast(float, hb & 0x7f800000u);
  float q = 0.0f;
  if (p2 != 0.0f) q = rintf((y - hi) * 32768.0f / p2);
  q = fminf(fmaxf(q, -127.0f), 127.0f);
  l = static_cast<int8_t>(static_cast<int>(q));
}

__host__ __device__ __forceinline__ float bf16c_decode(uint16_t h, int8_t l) {
  const uint32_t hb = static_cast<uint32_t>(h) << 16;
  const float hi = __builtin_bit_cast(float, hb);
  const float p2 = __builtin_bit_cast(float, hb & 0x7f800000u);
  return hi + static_cast<float>(l) * (p2 * (1.0f / 32768.0f));
}

// Byte offsets are formed in 32 bits from the element index (the host keeps
// 3 x values below 2^32), so every access is a wave-uniform base plus a
// 32-bit lane offset: value i of block i / 64, lane l = i % 64, has its bf16
// word at 3 (i - l) + 2 l = 3 i - l and its residual at 3 i - 2 l + 128.
template <>
struct Store<Bf16c> {
  using C = float;
  static __device__ __forceinline__ uint64_t off_hi(int64_t i) {
    const uint32_t u = (uint32_t)i;
    return (uint64_t)(3u * u - (u & 63u));
  }
  static __device__ __forceinline__ uint64_t off_lo(int64_t i) {
    const uint32_t u = (uint32_t)i;
    return (uint64_t)(3u * u - 2u * (u & 63u) + 128u);
  }
  static __device__ __forceinline__ const unsigned char* bytes(const Bf16c* p) {
    return reinterpret_cast<const unsigned char*>(p);
  }
  static __device__ __forceinline__ unsigned char* bytes(Bf16c* p) { return reinterpret_cast<unsigned char*>(p); }
  static __device__ __forceinline__ C ld(const Bf16c* p, int64_t i, C inv) {
    return bf16c_decode(*gp(reinterpret_cast<const uint16_t*>(bytes(p) + off_hi(i))),
                        *gp(reinterpret_cast<const int8_t*>(bytes(p) + off_lo(i)))) *
           inv;
  }
  static __device__ __forceinline__ void st(Bf16c* p, int64_t i, C v, C scale) {
    uint16_t h;
    int8_t r;
    bf16c_encode(v * scale, h, r);
    *gp(reinterpret_cast<uint16_t*>(bytes(p) + off_hi(i))) = h;
    *gp(reinterpret_cast<int8_t*>(bytes(p) + off_lo(i))) = r;
  }
  static __device__ __forceinline__ C ld_nt(const Bf16c* p, int64_t i, C inv) {
    return bf16c_decode(__builtin_nontemporal_load(gp(reinterpret_cast<const uint16_t*>(bytes(p) + off_hi(i)))),
                        __builtin_nontemporal_load(gp(reinterpret_cast<const int8_t*>(bytes(p) + off_lo(i))))) *
           inv;
  }
  static __device__ __forceinline__ void st_nt(Bf16c* p, int64_t i, C v, C scale) {
    uint16_t h;
    int8_t r;
    bf16c_encode(v * scale, h, r);
    __builtin_nontemporal_store(h, gp(reinterpret_cast<uint16_t*>(bytes(p) + off_hi(i))));
    __builtin_nontemporal_store(r, gp(reinterpret_cast<int8_t*>(bytes(p) + off_lo(i))));
  }
};

// Loads and stores at a small unsigned element offset from a wave-uniform
// base pointer: the offset is zero-extended from 32 bits, which lets the
// compiler address with the SGPR base + 32-bit VGPR offset form of the
// global instructions (no per-lane 64-bit address arithmetic).
template <typename S>
__device__ __forceinline__ typename Store<S>::C ld32(const S* p, uint32_t i, typename Store<S>::C inv) {
  return Store<S>::ld(p, (int64_t)(uint64_t)i, inv);
}
template <typename S>
__device__ __forceinline__ typename Store<S>::C ld32_nt(const S* p, uint32_t i, typename Store<S>::C inv) {
  return Store<S>::ld_nt(p, (int64_t)(uint64_t)i, inv);
}
template <typename S>
__device__ __forceinline__ void st32(S* p, uint32_t i, typename Store<S>::C v, typename Store<S>::C scale) {
  Store<S>::st(p, (int64_t)(uint64_t)i, v, scale);
}
template <typename S>
__device__ __forceinline__ void st32_nt(S* p, uint32_t i, typename Store<S>::C v, typename Store<S>::C scale) {
  Store<S>::st_nt(p, (int64_t)(uint64_t)i, v, scale);
}

template <typename S>
struct Fields {
  S* e[3];
  S* h[3];
};

template <typename C>
struct Scales {
  C inv_e, inv_h;  // stored -> physical
  C s_e, s_h;      // physical -> stored
};

struct Dims {
  int nx, ny, nz;  // cells
  int64_t sx, sy;  // strides of i and j (k stride is 1 in 3-D; unused in 2-D)
};

// Work decomposition of the 3-D marching kernels: a 1-D grid of nkb * njb *
// nch blocks; block (kb, jb, ch) owns k in [64 kb, 64 kb + 64), j in
// [ty jb, ty jb + ty) and marches i through [chunk ch, chunk (ch + 1)).
// With xcd != 0 the hardware block id is remapped so that each of the 8 XCDs
// (which receive hardware ids round-robin) owns a contiguous run of
// logical tiles (k fastest, then j), keeping the j/k halo rows of
// neighbouring tiles in the same XCD's L2.
struct Tiling {
  int nkb, njb, nch;
  int chunk;
  int xcd;
};

__device__ __forceinline__ void tile_coords(const Tiling& T, int& kb, int& jb, int& ch) {
  const int nblk = T.nkb * T.njb * T.nch;
  int id = blockIdx.x;
  if (T.xcd) {
    const int q = nblk / 8, r = nblk % 8;
    const int x = id % 8, n = id / 8;
    id = x * q + min(x, r) + n;
  }
  kb = id % T.nkb;
  id /= T.nkb;
  jb = id % T.njb;
  ch = id / T.njb;
}

// One PML face for one pass (E or H).  Component A is E_A/H_A of the face's
// (a, b, c) frame, B is E_B/H_B (pml_corrector.cu:50-270).
template <typename C>
struct PmlFace {
  int axis;
  int c0, c1;          // corrected nodes along the main axis
  int a0, a1, b0, b1;  // lateral task (extended by one per component)
  int gstart, nstart, off_c;
  const C* ca;
  const C* cb;
  C* psi_a;
  const C* cpsi_a;
  C* psi_b;
  const C* cpsi_b;
  int64_t sa1, sa2, sb1, sb2;  // shape[1], shape[2] of the two psi arrays
  int ta, tb;  // the coefficient (0..8) of a node's table row equal to every
               // c_psi value of psi_a / psi_b, or -1 (read the c_psi array)
};

template <typename C>
struct PmlPass {
  int n;
  PmlFace<C> f[kMaxPml];
};

// The z faces of a pass applied at the end of each Yee tile (at most two,
// in registration order), or n = 0.
template <typename C>
struct PmlZTail {
  int n;
  PmlFace<C> f[2];
};

// A pass's coefficient table and per-node ids (uint8, or uint16 when wide),
// for the CPML corrections that take c_psi from the node's row
// (PmlFace::ta / tb).
template <typename C>
struct RowRef {
  const C* tab;
  const void* ids;
  int wide;
};

// The z faces of a pass as node boxes (k fastest) for k_pml_z: face q
// covers i in [i0, i0 + ni), j in [j0, j0 + nj), k in [k0, k0 + nk), its
// nodes numbered from start[q]; start[n] is the total.
struct PmlZBox {
  int n;
  int i0[2], j0[2], k0[2];
  int nj[2], nk[2];
  int64_t start[3];
};

// Dense per-cell coefficients of one pass: array 3 c + m holds coefficient
// m (self, curl a, curl b) of component c at every cell, in the padded
// layout of the fields (the reference's per-field coefficient arrays,
// fdtd_coefficient.cuh:149-167).  Used when a scene's coefficient rows are
// too many to deduplicate into a table; a[0] == nullptr means table mode.
template <typename C>
struct Dense9 {
  const C* a[9];
};

// A z = const TF/SF correction left to the Yee tiles (XF_TILE_TFSF_ZTAIL):
// the written nodes at k that no other face of the same component also
// writes, [i0, i1) x [j0, j1); the incident value is taken at kinc (k, or
// the node beyond the face on its negative side).
template <typename C>
struct TfsfZJob {
  int comp, dual;
  int k, kinc;
  int i0, i1, j0, j1;
  C coef;
};

// The z = const TF/SF corrections of a pass applied after each tile's
// march, with what the incident interpolation needs (field names as in
// TfsfPass, so tfsf_inc reads either).
template <typename C>
struct TfsfZTail {
  int n;
  int t_arg;
  const int* t_dev;
  TfsfZJob<C> z[4];
  const C* proj_int[3];
  const C* proj_half[3];
  const C* inc;
  int64_t n_aux;
  int g[3];
  C tr[3];
};

// Everything one 3-D Yee pass (H or E) of one grid needs: the kernels take
// it as a kernel argument (one grid) or from an array in device memory
// (a batch of grids, one launch).
template <typename S>
struct YeePass {
  using C = typename Store<S>::C;
  Fields<S> f;
  const void* ids;              // uint8_t or uint16_t per cell (padded layout)
  const C* tab;                 // 9 coefficients per id
  int ntab;
  const unsigned char* eflag;   // E pass: table rows left to the ADE kernel
  const C* dense[9];            // dense coefficients (padded layout), or null
  Dims d;
  Scales<C> sc;
  int nkb, njb, chunk, nch;     // tiling of this grid
  PmlPass<C> pml;
  PmlZTail<C> zt;               // z faces applied after each tile's march
  TfsfZTail<C> tz;              // z = const TF/SF faces, likewise (n = 0: none)
};

// A TFSF face correction of one component (one reference launch of
// __tFSFcorrect, tfsf_corrector.cu:148-153) as a written-cell box.
template <typename C>
struct TfsfJob {
  int comp, face;
  int o[3], n[3];  // written box origin and extent
  int64_t start;   // first flat thread index of the job
  C coef;          // c{a|b}{axis} * coefficient_flag
  int edge_axis;   // -1: a thread per written node; else only the nodes at
                   // the two ends of this axis (the rest is a TfsfZJob's)
};

template <typename C>
struct TfsfPass {
  int njob;
  int is_e;
  TfsfJob<C> job[kMaxTfsfJobs];
  int64_t total;
  const C* proj_int[3];
  const C* proj_half[3];
  const C* inc;  // e_inc for the H pass, h_inc for the E pass
  int64_t n_aux;
  int g[3];  // total_task start
  C tr[3];   // transform_e (H pass) or transform_h (E pass)
};

struct NfFace {
  int o[3], n[3];
  int64_t start;  // first flat thread index
  int64_t off;    // offset of the face in a frequency block (cells)
};

struct NfPass {
  NfFace face[6];
  int64_t total;     // surface cells over the 6 faces
  int64_t block;     // cells of one frequency block (sum of face sizes)
  int nfreq;
  int64_t nt;
  double2* cur;      // [freq][face cells][4] -> laid out [freq][face][comp][cell]
  const double2* te; // [freq][nt]
  const double2* th;
};

// NF2FF time domain (one box).  A face's surface cells are cut into tiles
// of NT_TILE x NT_TILE cells over its two tangential axes (u < v in x, y, z
// order); one block per tile sums the tile's contributions per time bin
// (k_nftd_tile), one wave per bin then adds the tiles' partial sums in a
// fixed order (k_nftd_reduce).
constexpr int NT_TILE = 16;
constexpr int NT_MAXSPAN = 256;  // time bins one tile may touch in a step (LDS: 52 KiB)

struct NtFace {
  int o[3];         // task origin
  int nu, nv;       // extents along the tangential axes u < v
  int tu, tv;       // tiles along u and v
  int blk0;         // first tile of the face
  int bins;         // bin slots of the face's reduce (>= bins touched in a step)
  int rblk0, nrblk; // the face's reduce blocks
  int64_t prev_off; // first prev value of the face: [ea, eb, ha, hb][n_a][n_b]
  int nb;           // extent along b (prev row length)
  double rd_min, rd_max;  // range of r . r_unit over the face's cells
};

struct NtPass {
  NtFace face[6];
  int nblk, nrblk;
  int maxspan;      // bins of one tile's partial
  int64_t n_pot;
  double u[3];
  double dmax_e, dmax_h, c_dt;
  const double* e_node[3];
  const double* h_node[3];
  const double* e_size[3];
  double* pot;      // [face][wa, wb, ua, ub][n_pot]
  double* prev;     // per face at prev_off
  double* part;     // [tile][wa, wb, ua, ub][maxspan]
  int* hdr;         // [tile][bmin_e, span_e, bmin_h, span_h]
};

// Point sources: node index (padded layout) and component of each; the
// values are a [source][step] table of doubles.
struct PointDev {
  int n;
  int64_t nt;
  const int* comp;
  const int64_t* idx;
  const double* value;
};

// A movie monitor's box and its frame buffer (doubles, frames x box).
struct MovieDev {
  int field;           // 0..2 E, 3..5 H
  int o[3], n[3];      // box origin and extent
  int64_t interval, frames, box;
  double* data;
};

// The learned layer (xf_learned_boundary): the faces in the mask, in face
// order, with their column ranges; columns are numbered face by face,
// component E_a then E_b, then a-index, then b-index.
struct LbFace {
  int f, c, a, b, positive;
  int tr;              // y faces: threads walk a (= z) fastest and the device
                       // history is stored [b][a], so both stream along z
  int64_t hist_off;    // first history value of the face
  int64_t na1, nb1;    // n_a + 1, n_b + 1 (history plane)
  int64_t start[2];    // first column of E_a, E_b
  int p1[2], q1[2];    // last a-, b-index of E_a, E_b (first is inset)
};

// The MLP weights as the matrix-core predictor's per-lane operands, built
// on the host from the fp16 weights.  For lane l of a wave, g = l / 32:
// a1 holds W1[l % 32][8g + j] (zero for inputs 12..15); a2[p] holds
// W2[l % 32][16p + 8(j / 4) + 4g + j % 4], the hidden unit the layer-1
// result row of that k-slot carries; c1, c2 and w3 hold b1, b2 and w3 of
// unit 8(i / 4) + 4g + i % 4, the row of accumulator element i.
struct LbPack {
  uint16_t a1[64][8];     // fp16 bits
  uint16_t a2[2][64][8];  // fp16 bits
  float c1[16][64];
  float c2[16][64];
  float w3[16][64];
  float b3;
};

struct LbDev {
  int mode, faces, inset, nface;
  float mur_k;
  const LbPack* pack;  // APPLY: the weights as MFMA operands
  float* hist;
  float* rec;          // RECORD: [slot][column][13]
  int rec_every;
  int64_t rec_slots;
  int64_t total;       // columns
  int n[3];
  LbFace face[6];
  // APPLY: the columns whose layer-0 node lies on two faces (edges), for the
  // edge phase and their history roll; the other columns roll their history
  // in the matrix-core phase
  const uint32_t* edge;
  int64_t n_edge;
};

// One grid's TFSF pass, NF2FF box, point sources or learned layer inside a
// batch of grids: the batched kernels read these from device memory, one per
// grid (or box).
template <typename S>
struct TfsfItem {
  using C = typename Store<S>::C;
  Fields<S> f;
  Scales<C> sc;
  Dims d;
  TfsfPass<C> P;
};

template <typename S>
struct NfItem {
  using C = typename Store<S>::C;
  Fields<S> f;
  Scales<C> sc;
  Dims d;
  NfPass P;
};

template <typename S>
struct PmlZItem {
  using C = typename Store<S>::C;
  Fields<S> f;
  Scales<C> sc;
  Dims d;
  PmlPass<C> P;  // the z faces
  PmlZBox B;
  RowRef<C> R;
};

template <typename S>
struct PointItem {
  using C = typename Store<S>::C;
  Fields<S> f;
  Scales<C> sc;
  PointDev P;
};

template <typename S>
struct LbItem {
  using C = typename Store<S>::C;
  Fields<S> f;
  Scales<C> sc;
  Dims d;
  LbDev L;
};

}  // namespace xf

#endif  // XF_KERNELS_H_
