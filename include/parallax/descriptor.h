/*
 * parallax/descriptor.h -- the primitive descriptor: the payload that
 * parallax_kernel_load / parallax_kernel_register take in place of SPIR-V.
 *
 * The reference registers compiler-emitted SPIR-V under each funnel key
 * (reference runtime.h:96-103, src/kernel/cache.cpp:71-104).  The MI355X engine ships
 * its kernels precompiled for gfx950, so what a registrar hands over is a small,
 * versioned word array naming which kernel family, element type and predicate / map
 * functor to run.  The (const unsigned*, size_t words) signature is unchanged.
 *
 * Layout: PX_DESC_WORDS 32-bit words, see struct px_desc.  Constants are stored as the
 * little-endian bytes of the element type (low 4 bytes for 32-bit types).
 */
#ifndef PARALLAX_DESCRIPTOR_H
#define PARALLAX_DESCRIPTOR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PX_DESC_MAGIC 0x50584D49u /* "IMXP" little-endian */
#define PX_DESC_VERSION 1u
#define PX_DESC_WORDS 12u

/* Element types. */
enum px_dtype {
  PX_I32 = 0,
  PX_U32 = 1,
  PX_I64 = 2,
  PX_U64 = 3,
  PX_F32 = 4,
  PX_F64 = 5,
  PX_NUM_DTYPES = 6
};

/* Kernel families.  The scan family keeps the reference's three roles
 * (":scan", ":add", ":shift"); the engine's single-pass scan runs from the SCAN
 * handle and only validates the other two. */
enum px_op {
  PX_OP_NONE = 0,
  PX_OP_REDUCE = 1,     /* sum, identity 0 */
  PX_OP_SCAN = 2,       /* inclusive scan tile kernel (":scan") */
  PX_OP_SCAN_ADD = 3,   /* block-offset add (":add"), folded into SCAN */
  PX_OP_SCAN_SHIFT = 4, /* exclusive finalize (":shift"), folded into SCAN */
  PX_OP_SORT = 5,       /* ascending sort by operator< */
  PX_OP_FLAGS = 6,      /* compaction predicate (":flags"), pred in w[5] */
  PX_OP_SCATTER = 7,    /* compaction scatter (":scatter"), mode in w[4] */
  PX_OP_MAP = 8,        /* in-place map x = f(x) (launch / launch_with_captures) */
  PX_OP_TRANSFORM = 9,  /* out[i] = f(in[i]) with in dtype w[3], out dtype w[4] */
  PX_OP_ARGMINMAX = 10, /* search family */
  PX_OP_FIND = 11,
  PX_OP_MISMATCH = 12,
  PX_OP_CODE = 13       /* a gfx950 code object carried in the payload (below) */
};

/* Code-object payload (op PX_OP_CODE), for functors no built-in map describes (a
 * compiler registers the gfx950 code object it emitted for the funnel's lambda):
 *   words 0..11  the descriptor header: dtype = input element type, aux = output element
 *                type, fn = px_code_kind, fn2 = length in bytes of the kernel symbol
 *                name, c0 = size in bytes of the code object, c1 = 0;
 *   then         the symbol name (no terminating NUL), zero-padded to a word boundary;
 *   then         the code object (an HSACO for gfx950), zero-padded to a word boundary.
 * The kernel is launched with 256-thread blocks and ceil(count / 256) blocks, one element
 * per thread, on the runtime's stream, with these extern "C" signatures:
 *   PX_CODE_MAP        (launch, launch_with_captures: in place)
 *     void k(void* data, uint64_t count, const void* captures, uint64_t capture_size,
 *            uint64_t host_base, uint64_t dev_base);
 *   PX_CODE_TRANSFORM  (launch_transform2 / _captures: in -> out)
 *     void k(const void* in, void* out, uint64_t count, const void* captures,
 *            uint64_t capture_size, uint64_t host_base, uint64_t dev_base);
 * `captures` is a device copy of the closure bytes (null when there are none);
 * host_base / dev_base relocate a captured arena pointer p: dev_base + (p - host_base)
 * (both 0 without an arena). */
enum px_code_kind {
  PX_CODE_MAP = 0,
  PX_CODE_TRANSFORM = 1
};

/* Compaction scatter modes (w[4] of a PX_OP_SCATTER descriptor). */
enum px_scatter_mode {
  PX_SCATTER_COMPACT = 0,          /* copy_if / remove_if / unique: kept elements only */
  PX_SCATTER_PARTITION = 1,        /* every element, in std::partition's order (the partition
                                      funnel's fallback): kept first, num_true returned */
  PX_SCATTER_STABLE_PARTITION = 2  /* every element: kept first, the rest after num_true,
                                      both in input order (std::stable_partition) */
};

/* Predicates (w[5] of a PX_OP_FLAGS / PX_OP_FIND / count descriptor).  `c0`, `c1`
 * are the constants.  w[6] bit 0 negates the predicate (remove_if, find_if_not). */
enum px_pred {
  PX_PRED_TRUE = 0,
  PX_PRED_EQ = 1,        /* x == c0 */
  PX_PRED_NE = 2,        /* x != c0 */
  PX_PRED_LT = 3,        /* x <  c0 */
  PX_PRED_LE = 4,        /* x <= c0 */
  PX_PRED_GT = 5,        /* x >  c0 */
  PX_PRED_GE = 6,        /* x >= c0 */
  PX_PRED_IN_RANGE = 7,  /* c0 <= x && x < c1 */
  PX_PRED_MOD_EQ = 8,    /* x % c0 == c1 (C++ truncated remainder; fmod for floats) */
  PX_PRED_NONZERO = 9,   /* x != 0 */
  PX_PRED_BITS_ANY = 10, /* (x & c0) != 0, integer types */
  PX_PRED_UNIQUE = 11,   /* i == 0 || x[i] != x[i-1] (std::unique keeps the first) */
  PX_NUM_PREDS = 12
};
#define PX_FLAG_NEGATE 1u

/* Map functors (w[5] of PX_OP_MAP / PX_OP_TRANSFORM).  `a`, `b` come from the
 * capture bytes when capture_size covers them (the closure's members, in order, of
 * the element type), else from c0 / c1. */
enum px_mapop {
  PX_MAP_IDENTITY = 0, /* y = x (transform: converting copy) */
  PX_MAP_FILL = 1,     /* y = a */
  PX_MAP_ADD = 2,      /* y = x + a */
  PX_MAP_MUL = 3,      /* y = x * a */
  PX_MAP_AXPB = 4,     /* y = a * x + b */
  PX_MAP_NEG = 5,      /* y = -x */
  PX_MAP_ABS = 6,      /* y = |x| */
  PX_MAP_SQUARE = 7,   /* y = x * x */
  PX_MAP_PRED = 8,     /* y = pred(x) ? 1 : 0 (count_if ":pred"); pred in w[7] */
  PX_MAP_GATHER = 9,   /* y = table[x]; capture = {const T* table (arena/host-mapped
                          pointer, relocated on the device), uint64 size} */
  PX_NUM_MAPOPS = 10
};

/* The descriptor as a struct (identical to the word layout). */
typedef struct px_desc {
  uint32_t magic;    /* w[0] PX_DESC_MAGIC */
  uint32_t version;  /* w[1] PX_DESC_VERSION */
  uint32_t op;       /* w[2] px_op */
  uint32_t dtype;    /* w[3] px_dtype (input type) */
  uint32_t aux;      /* w[4] out dtype (TRANSFORM) or scatter mode (SCATTER) */
  uint32_t fn;       /* w[5] px_pred or px_mapop */
  uint32_t flags;    /* w[6] PX_FLAG_* */
  uint32_t fn2;      /* w[7] predicate of PX_MAP_PRED */
  uint64_t c0;       /* w[8..9]  */
  uint64_t c1;       /* w[10..11] */
} px_desc;

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* PARALLAX_DESCRIPTOR_H */
