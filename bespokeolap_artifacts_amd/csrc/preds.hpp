// preds.hpp — device evaluation of bolap_pred conjunctions (dimension filters).
#pragma once

#include "common.hpp"

namespace bolap {

struct PredK {
    const void* col;
    int32_t dtype;
    int32_t op;
    int64_t lo, hi;
    uint64_t inl[4];
    const uint32_t* set;
    int64_t set_bits;
};

struct PredSet {
    int32_t n;
    PredK p[BOLAP_MAX_PREDS];
};

inline int make_predset(PredSet& ps, int32_t n, const bolap_pred* preds, int64_t rows,
                        const char* who) {
    BOLAP_ARG_CHECK(n >= 0 && n <= BOLAP_MAX_PREDS, "%s: %d predicates (max %d)", who, n,
                    BOLAP_MAX_PREDS);
    ps.n = n;
    for (int i = 0; i < n; ++i) {
        const bolap_pred& q = preds[i];
        BOLAP_ARG_CHECK(rows == 0 || q.col, "%s: predicate %d has no column", who, i);
        BOLAP_ARG_CHECK(dtype_size(q.dtype) > 0, "%s: predicate %d dtype %d", who, i, q.dtype);
        BOLAP_ARG_CHECK(q.op == BOLAP_OP_RANGE || q.op == BOLAP_OP_IN_SET,
                        "%s: predicate %d op %d", who, i, q.op);
        BOLAP_ARG_CHECK(q.op != BOLAP_OP_IN_SET || !q.set || q.set_bits >= 0,
                        "%s: predicate %d set_bits < 0", who, i);
        PredK& k = ps.p[i];
        k.col = q.col;
        k.dtype = q.dtype;
        k.op = q.op;
        k.lo = q.lo;
        k.hi = q.hi;
        for (int w = 0; w < 4; ++w) k.inl[w] = q.inline_set[w];
        k.set = q.set;
        k.set_bits = q.set_bits;
    }
    return BOLAP_OK;
}

BOLAP_DEV int64_t load_any(const void* col, int32_t dtype, int64_t row) {
    switch (dtype) {  // uniform across the wave
        case BOLAP_U8: return static_cast<const uint8_t*>(col)[row];
        case BOLAP_I16: return static_cast<const int16_t*>(col)[row];
        case BOLAP_I32: return static_cast<const int32_t*>(col)[row];
        default: return static_cast<const int64_t*>(col)[row];
    }
}

BOLAP_DEV bool pred_holds(const PredK& p, int64_t v) {
    if (p.op == BOLAP_OP_RANGE) return v >= p.lo && v <= p.hi;
    if (p.set) {
        if (v < 0 || v >= p.set_bits) return false;
        return (p.set[v >> 5] >> (v & 31)) & 1u;
    }
    if (v < 0 || v > 255) return false;
    const int w = int(v >> 6);
    const uint64_t word = w == 0 ? p.inl[0] : w == 1 ? p.inl[1] : w == 2 ? p.inl[2] : p.inl[3];
    return (word >> (v & 63)) & 1u;
}

// R consecutive rows of one column (rows unit*R .. unit*R+R-1) with one
// vector load; the column must be aligned to R elements.
template <int R, typename T>
BOLAP_DEV void load_rows(const void* col, int64_t unit, int64_t (&v)[R]) {
    const VecT<T, R> x = load_unit<T, R>(static_cast<const T*>(col), unit);
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = int64_t(x.v[r]);
}

// Conjunction over R consecutive rows: bit r of the result is row unit*R+r.
// Every predicate is evaluated for every row (no short circuit), so the
// loads of all predicate columns are issued together.
template <int R>
BOLAP_DEV uint32_t eval_preds_unit(const PredSet& ps, int64_t unit) {
    uint32_t m = (1u << R) - 1u;
    for (int i = 0; i < ps.n; ++i) {
        const PredK& p = ps.p[i];
        int64_t v[R];
        switch (p.dtype) {  // uniform across the wave
            case BOLAP_U8: load_rows<R, uint8_t>(p.col, unit, v); break;
            case BOLAP_I16: load_rows<R, int16_t>(p.col, unit, v); break;
            case BOLAP_I32: load_rows<R, int32_t>(p.col, unit, v); break;
            default: load_rows<R, int64_t>(p.col, unit, v); break;
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (!pred_holds(p, v[r])) m &= ~(1u << r);
    }
    return m;
}

// True when every predicate column is aligned for eval_preds_unit<R>.
inline bool preds_aligned(const PredSet& ps, int R) {
    for (int i = 0; i < ps.n; ++i)
        if (!aligned(ps.p[i].col, size_t(R) * size_t(dtype_size(ps.p[i].dtype)))) return false;
    return true;
}

BOLAP_DEV bool eval_pred(const PredK& p, int64_t row) {
    return pred_holds(p, load_any(p.col, p.dtype, row));
}

BOLAP_DEV bool eval_preds(const PredSet& ps, int64_t row) {
    bool ok = true;
    for (int i = 0; i < ps.n; ++i) ok = ok && eval_pred(ps.p[i], row);
    return ok;
}

}  // namespace bolap
