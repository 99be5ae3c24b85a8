// Internal declarations shared by the HIP translation units of libtimeevolver_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/timeevolver.h"

namespace te {

constexpr int kMaxKrylov = 64;      // Krylov dimension cap: one wavefront holds a row of Q
constexpr int kBlock = 256;         // threads per block of the streaming kernels

extern int g_log_level;
void log_msg(int level, const char* fmt, ...);

// Sets ctx->err and returns code.
int fail(te_ctx* ctx, int code, const char* fmt, ...);

#define TE_HIP(ctx, call)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::te::fail((ctx), TE_ERR_HIP, "%s failed: %s (%s:%d)", #call,          \
                              hipGetErrorString(e_), __FILE__, __LINE__);                  \
    } while (0)

struct SpmvShape {
    int kind;             // 0: LPR lanes per row; 1: row blocks of ~kStreamChunk entries staged in LDS
    int lanes_per_row;    // lanes of a wavefront cooperating on one row (1..64, power of 2)
    int grid;             // blocks of the launch
};

// Row-block ("CSR-stream") SpMV: block b holds the rows r with ptr[r] + r in
// [b * kStreamChunk, (b+1) * kStreamChunk); its entries (at most kStreamChunk + longest row - 1)
// are multiplied into LDS with every load issued up front, then reduced row by row.
constexpr int kStreamChunk = 2048;
constexpr int kStreamCap = kStreamChunk + 256;     // LDS entries per block (rows up to 257 long)

// Column split: when x is larger than the last-level cache (256 MB MALL on MI355X), the matrix
// is also held as kMaxParts or fewer column slices, each gathering from a slice of x that
// stays cache-resident during its pass; partial row sums are carried between passes.
constexpr int kMaxParts = 8;
struct CsrPart {
    int* ptr = nullptr;       // n_rows + 1
    int* col = nullptr;
    double2* val = nullptr;
    int* row_block = nullptr; // n_row_blocks + 1
    int n_row_blocks = 0;
    int64_t nnz = 0;
};

// Value-dictionary copy of a matrix (lossless): the off-diagonal entries keep their CSR order
// with 32-bit columns and 16-bit indices into a table of the distinct values (compared bit for
// bit), the diagonal is a dense vector. It streams 6 bytes per off-diagonal entry instead of 20.
// Row blocks as in the stream kernel, with a smaller chunk so that the products (kVdCap entries)
// and the table (kVdMaxValues entries) fit 40 KB of LDS with the block's flags: four blocks
// per CU (160 KB).
constexpr int kVdChunk = 1280;
constexpr int kVdCap = 1536;        // rows up to kVdCap - kVdChunk + 1 = 257 off-diagonal entries
constexpr int kVdMaxValues = 1008;
constexpr int kVdPad = 256;         // zero entries after col/vidx: clamped loads stay in bounds
struct VdCsr {
    int* ptr = nullptr;             // n_rows + 1, off-diagonal entries
    int* col = nullptr;
    unsigned short* vidx = nullptr;
    double2* vtab = nullptr;        // n_values distinct off-diagonal values
    int n_values = 0;
    double2* diag = nullptr;        // n_rows, zero where the row stores no diagonal entry
    int2* blocks = nullptr;         // n_row_blocks + 1: (first row, its first entry)
    int n_row_blocks = 0;
    int grid = 0;
    int64_t nnz = 0;
    // Optional locality order of the rows and columns: the copy holds P A P^T with
    // perm[new] = old, iperm[old] = new. Vectors the Lanczos kernels see are then in that order.
    int* perm = nullptr;
    int* iperm = nullptr;
    // Sliced-ELL form of the same entries (when every row has at most 255 off-diagonal entries):
    // slices of kSlice consecutive rows. Shared slots: the leading and trailing entry positions
    // at which every row of a full slice has an entry with the same column offset from its row
    // and the same value index (transitions that change only modes the slice's rows share) are
    // stored once per slice as records (column - row, value index): the first kSellFix in
    // sell_ufix[s * kSellFix + q], the rest in sell_uni from the slice's overflow offset. The
    // entries between them (the middle of each row) are stored slot-major, slot j of row
    // s*kSlice + l at word meta[s].x + j*kSlice + l, so a wavefront reads every slot of its 64
    // rows with one coalesced load. Each row's middle entries keep their column order; they are
    // dealt to slots so that the columns of one slot (over the slice's rows) lie within
    // kSellSpan of each other: slot j starts at column sell_base[meta[s].x / kSlice + j], and a
    // word holds the 16-bit offset from it and, above, the 16-bit value index (kSellPadIdx: no
    // entry, not added; the offset is then 0, a column of the slot). A row sums diag, leading
    // shared, middle, trailing shared: its column order. sell_meta[s] = (first word, first
    // overflow record, leading shared records | shared records << 16, 0); sell_meta[n_slices] =
    // (words, overflow records, 0, 0). Optionally (TE_SELL_SIGMA=1) rows are ordered by
    // decreasing length within windows of kSellWindow (part of perm): less padding, worse gathers.
    bool sell = false;
    int4* sell_meta = nullptr;          // n_slices + 1
    unsigned* sell_word = nullptr;      // sell_entries + kSellGroup * kSlice (zero padding)
    int* sell_base = nullptr;           // sell_entries / kSlice + kSellGroup (zero padding)
    int2* sell_ufix = nullptr;          // n_slices * kSellFix, or null without shared slots
    int2* sell_uni = nullptr;           // overflow records (at least one)
    int n_slices = 0;
    int64_t sell_entries = 0;
    int64_t sell_shared = 0;            // shared records
    int64_t sell_overflow = 0;          // of them in sell_uni
    int sell_grid = 0;
    int sell_chunk = 0;                 // SellView::chunk
    // The real parts of diag when every imaginary part is +0.0 (bit pattern 0), else nullptr:
    // the sliced-ELL kernels then stream 8 bytes a row for the diagonal and multiply by
    // (re, +0.0) held in registers, the same arithmetic on the same operands.
    double* diag_re = nullptr;
};

constexpr int kSlice = 64;
constexpr int kSellGroup = 4;       // slots a lane loads (and gathers) before using any
constexpr int kSellWindow = 512;
constexpr int kSellSpan = 1 << 16;  // columns one slot of a slice may span (16-bit offsets)
constexpr unsigned kSellPadIdx = 0xffffu;   // value index of a padding slot
constexpr int kSellFix = 12;        // shared records per slice at fixed places (sell_ufix)
static_assert(kVdMaxValues < (int)kSellPadIdx, "value indices must stay below the padding index");

// What the sliced-ELL kernels read.
struct SellView {
    const int4* meta;
    const unsigned* word;
    const int* base;
    const double2* vtab;
    int n_values;
    const double2* diag;
    int n_slices;
    int64_t n;
    int chunk;      // 0: XCD x sweeps the x-th eighth of the slices; > 0: chunks of `chunk` slices
                    // dealt round-robin to the XCDs, so all XCDs sweep one front together
    const double* diag_re;   // real diagonal (VdCsr::diag_re) or nullptr: read diag
    const int2* ufix;        // shared records (VdCsr::sell_ufix) or nullptr
    const int2* uni;
};

// What the kernels read of a VdCsr.
struct VdView {
    const int* ptr;
    const int* col;
    const unsigned short* vidx;
    const double2* vtab;
    int n_values;
    const double2* diag;
    const int2* blocks;
    int n_row_blocks;
};

}  // namespace te

struct te_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 256;
    int max_lds = 65536;        // bytes of LDS one workgroup may use
    std::string err;
};

struct te_matrix {
    te_ctx* ctx = nullptr;
    int64_t n_rows = 0, n_cols = 0, nnz = 0;
    int* row_ptr = nullptr;     // device, n_rows + 1
    int* col = nullptr;         // device, nnz
    double2* val = nullptr;     // device, nnz complex128
    double norm1 = 0.0;
    bool norm1_valid = false;
    bool diagonal = false;      // every entry on the diagonal (observables of n_k type)
    double* diag_re = nullptr;  // device, n_rows: Re of the diagonal when `diagonal`
    te::SpmvShape shape{0, 8, 1024};
    int* row_block = nullptr;   // device, n_row_blocks + 1 first rows of the stream row blocks
    int n_row_blocks = 0;
    int max_row_len = 0;
    int n_parts = 0;            // > 1: column-split copy in parts[] used by the stream kernels
    te::CsrPart parts[te::kMaxParts];
    double2* acc = nullptr;     // n_rows partial row sums between the passes of a split SpMV
    bool use_vd = false;        // the stream kernels read the value-dictionary copy
    te::VdCsr vd;
};

namespace te {

// Grid of the row-range streaming kernels (the adjoint's passes, the row-block SpMV) for
// work_items rows: a multiple of 8 (one share per XCD), at most TE_STREAM_BLOCKS_PER_CU
// (default 4) blocks per CU.
int stream_grid(const te_ctx* ctx, int64_t work_items);
// Grid of the Lanczos projection (TE_PROJECT_BLOCKS_PER_CU, default 2, blocks per CU).
int project_grid(const te_ctx* ctx, int64_t rows);
// Grid of the evolver's combine (TE_COMBINE_BLOCKS_PER_CU, default 2, blocks per CU).
int combine_grid(const te_ctx* ctx, int64_t rows);
// SpMV launch shape for a matrix (lanes per row from the mean row length, grid from the CUs).
SpmvShape choose_shape(const te_ctx* ctx, int64_t n_rows, int64_t nnz);
// Derives the row blocks of the stream kernel on the device and picks the kernel (call after
// row_ptr is on the device). perm (device, n_rows entries, new -> old), when given, is the row
// and column order of the value-dictionary copy the Lanczos kernels read; it is copied.
int prepare_spmv(te_ctx* ctx, te_matrix* A, const int* perm = nullptr);
// The sliced-ELL view of A's value-dictionary copy (valid when A->use_vd && A->vd.sell).
SellView sell_view_of(const te_matrix* A);
// Vector gathers between the reference order and a matrix's locality order.
// dst[i] = src[idx[i]] for i < n.
void launch_gather(const double2* src, const int* idx, double2* dst, int64_t n, hipStream_t s);
void launch_gather_re(const double* src, const int* idx, double* dst, int64_t n, hipStream_t s);
void free_parts(te_matrix* A);
void free_vd(te_matrix* A);

// ---- kernels launched across translation units (te_spmv.hip) ----
void launch_spmv(const te_matrix* A, double2 alpha, const double2* x, double2* y, hipStream_t s);
// Deterministic reductions with a last-block-done finish. partials: >= grid entries.
struct ReduceScratch {
    double2* partials = nullptr;   // grid * 128 doubles2 (a combine launch uses up to 128 slots a block)
    unsigned int* ticket = nullptr;
    double2* result = nullptr;     // 64 slots
    int grid = 0;
};
int reduce_scratch_alloc(te_ctx* ctx, ReduceScratch& rs, int grid);
void reduce_scratch_free(ReduceScratch& rs);
// result[slot] = sum conj(x) y  (device)
void launch_zdotc(const double2* x, const double2* y, int64_t n, ReduceScratch& rs, int slot, hipStream_t s);
// result[slot].x = sum |x|^2
void launch_nrm2sq(const double2* x, int64_t n, ReduceScratch& rs, int slot, hipStream_t s);

}  // namespace te
