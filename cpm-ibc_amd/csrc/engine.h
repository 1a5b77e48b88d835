// Internal types of the MI355X closest-point engine (libcpmgpu.so).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "refine.h"
#include "tuned.h"

namespace cpmgpu {

constexpr int kBlock = 256;          // 4 wave64 per workgroup
// Block map entries of a launch over several segments: the segment's index in
// the low kSegBits bits, the block's index within its segment above them.
constexpr int kSegBits = 11;
constexpr int kMaxSegs = 1 << kSegBits;
constexpr int kMaxNR = 8;            // right-hand sides per multi-RHS sweep

// Row flags (one byte per row).
enum : uint8_t { ROW_IDENTITY = 1, ROW_NODIAG = 2 };

enum : int { OP_ELL = 0, OP_TP = 1 };

// Device view of an operator in one precision.  Arrays are slot-major (entry
// s of row r at [s * D + r]) so that a wavefront's 64 rows read 64
// consecutive words per slot.
//
// OP_ELL: E and Loff are ELL matrices of widths ew and lw.
// OP_TP (operators built on a tube, p = 3): rows are in the tube's
// lexicographic order, so the four z-neighbours b + (i, j, 0..3) of a stencil
// line are consecutive columns; E row = sum over (i, j, k) of
// (wz[k] wy[j]) wx[i] x[tpc(i, j) + k] (the reference's slot order and
// product order), except for the slots flagged in the row's exception record,
// whose columns were moved by an interior boundary (and one optional extra
// entry): such a row lists the columns of all 64 slots in its record.
// L off-diagonals all equal loff (lcol -1 = absent), L_ii = ldu.
template <typename T>
struct OpView {
    int64_t D;
    int nD;
    int ew, lw;          // ELL widths of E and of L without its diagonal
    const int* ecol;
    const T* eval;
    const int* lcol;
    const T* lval;
    const T* ldiag;      // L_ii (0 where L has no diagonal)
    const uint8_t* flag;
    int fmt;             // OP_ELL / OP_TP
    const int* tpc;      // [16][D] first column of stencil line (i, j)
    const T* tpw;        // [12][D] wx[4], wy[4], wz[4]; a row's sign is folded into wx
    const int* exc;      // [D] exception record of the row, -1 = none
    const int* ecx;      // [nexc][64] column of every slot of an exception row
    const int* xcol;     // [nexc] extra column after the stencil, -1 = none
    const T* xval;       // [nexc]
    T loff, ldu;
    // LDS tiles of the tensor-product format (kBlock rows each): the sorted
    // distinct x rows the tile's stencil lines read, gathered into LDS once.
    const int* tln;      // [ntiles] rows in the tile's list, -1 = the tile reads x directly
    const int* tlo;      // [ntiles] offset of the tile's list in tlr
    const int* tlr;      // lists of x rows
    const int16_t* tpl;  // [16][D] position of line (i, j)'s first row in the tile's list
    const int16_t* tlx;  // [nexc][kExcPos] positions of an exception row's 64 columns and extra entry
    // Shared-stencil tiles (built beside the row tiles; gln, glr, glx are
    // their tln, tlr, tlx).  Rows whose E stencils are the same 4x4x4 block
    // of nodes are handled by one thread, which reads each x row of the block
    // from LDS once for all of them.  Tile t owns the item places t kBlock ..
    // (t + 1) kBlock - 1, one per thread (places past its items hold empty
    // items, kind 3), and the list places t kGrpLdsRows .. (t + 1)
    // kGrpLdsRows - 1 (gln[t] in use, the rest repeat the last): a tile's
    // data needs no lookup.  An item is up to kGrpRows rows with one stencil
    // (kind 0), one exception row (kind 1) or up to kGrpRows identity rows
    // (kind 2), listed in slots gofs[i] .. gofs[i + 1] - 1.
    int ngt;             // shared-stencil tiles (0: none)
    int nitem;           // item places, ngt kBlock
    int64_t nslot;
    const int* gofs;     // [nitem + 1] first slot | kind << kGrpKindShift
    const int* grow;     // [nslot] row of each slot
    const int* gpl;      // [8][nitem] kind 0: word h holds the positions of lines 2h and 2h + 1 (line
                         // (i, j) = 4 i + j) in the tile's list, low and high 16 bits; kind 1: word 0
                         // is the exception record; kind 2: words 0 .. kGrpRows - 1 are the rows (-1: none)
    const T* gw;         // [nslot][12] the slot's row's wx[4], wy[4], wz[4] (48 or 96 contiguous bytes)
    const int* gln;      // [ngt] rows in use in the tile's list
    const int* glr;      // [ngt][kGrpLdsRows] the tiles' lists of x rows
    const int16_t* glx;
};

constexpr double kDefaultUpdateDelta = 1e-2;   // FP32+refine: refine when the FP32 residual fell 100x
constexpr int kLdsRows = CPMGPU_LDS_ROWS;        // longest LDS tile list
constexpr int kLdsBytes = 32 * CPMGPU_LDS_ROWS;  // LDS of one E-gather workgroup (gfx950: 160 KB per CU)
constexpr int kExcPos = 72;      // positions per exception record (64 slots, extra entry, padding)
#ifndef CPMGPU_GROUP_ROWS
#define CPMGPU_GROUP_ROWS 2
#endif
#ifndef CPMGPU_GROUP_WAVES_WIDE
#define CPMGPU_GROUP_WAVES_WIDE 2
#endif
#ifndef CPMGPU_GROUP_LDS_ROWS
#define CPMGPU_GROUP_LDS_ROWS 1280
#endif
constexpr int kGrpRows = CPMGPU_GROUP_ROWS;   // rows per shared-stencil item
static_assert(kGrpRows >= 1 && kGrpRows <= 8, "an identity item's rows fit the 8 words of its line positions");
constexpr int kGrpLdsRows = CPMGPU_GROUP_LDS_ROWS;        // longest shared-stencil tile list
constexpr int kGrpLdsBytes = 32 * CPMGPU_GROUP_LDS_ROWS;  // LDS of one shared-stencil workgroup
// LDS rows of a shared-stencil tile for T x NR values per row (0: these
// tiles are not used for this value size).
template <typename T, int NR>
struct GrpTile {
    static constexpr int kRows = (kGrpLdsRows * NR * (int)sizeof(T) <= kGrpLdsBytes) ? kGrpLdsRows : 0;
};
constexpr int kGrpKindShift = 28;             // item kind above the slot offset in gofs
#ifndef CPMGPU_GROUP_MIN_TILES
#define CPMGPU_GROUP_MIN_TILES 256
#endif
constexpr int kGrpMinTiles = CPMGPU_GROUP_MIN_TILES;   // fewer shared-stencil tiles: row tiles only
constexpr int kGrpOfsMask = (1 << kGrpKindShift) - 1;

// One instance group of a batched launch: an operator, NR right-hand sides
// and the contiguous block range that covers its D rows.
template <typename T>
struct Seg {
    OpView<T> op;
    OpView<double> op64; // the same operator in FP64 (refinement steps of FP32 segments)
    double m, n;
    int64_t off;         // element offset of row 0 in every workspace vector (rows interleave NR rhs)
    int nblocks;
    int nr;              // rhs in this group (<= template NR; extra lanes idle)
    int precond;         // CPM_PRECOND_*
    int precond_iters;   // Jacobi sweeps
    T dinv, tinv;        // FP32: 1 / (m + ldu n) (tensor-product format) and 1 / (m - n), for P^-1
    KState* st;          // nr states
    double* part;        // [2 * NR][nblocks] partial sums of the last reducing kernel
};

void hip_check(hipError_t e, const char* what);
#define HIPCHK(x) ::cpmgpu::hip_check((x), #x)

struct Buffer {
    void* p = nullptr;
    size_t bytes = 0;
    Buffer() = default;
    explicit Buffer(size_t n);
    ~Buffer();
    Buffer(const Buffer&) = delete;
    Buffer& operator=(const Buffer&) = delete;
    Buffer(Buffer&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    Buffer& operator=(Buffer&& o) noexcept {
        if (this != &o) {
            if (p) (void)hipFree(p);
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    template <typename U> U* as() const { return static_cast<U*>(p); }
};

}  // namespace cpmgpu
