// plan.h -- the brick plan's fold of a solid's postfix program against a
// brick box, compiled for the device (k_plan) and for the host (the library
// checks its two folds against each other on random programs,
// mc3d_selftest_plan_fold).
//
// The fold keeps, per stack entry, {certainly false, certainly true, an
// expression whose code starts at output position n}; atoms whose
// conservative box misses the brick are false, boxes, bounds checks and
// spheres that contain it are true, and the operators simplify constants
// away.  plan_fold walks every instruction; plan_fold_nodes walks the same
// program as node records (PlanNode, one fetch per step) and jumps over
// subtrees whose atoms' union box misses the brick (on the device: every
// brick of the wave), pushing the constant that subtree has when all its
// atoms are false -- which is exactly what plan_fold computes for it, having
// emitted nothing for it -- so both produce the same folded code, the same
// root and the same overflow flag.
#pragma once

#include <math.h>
#include <stdint.h>

#include "eval.h"
#include "mc3d_internal.h"

namespace mc3d {

// Atom against a brick box: 0 certainly false, 1 certainly true, 2 evaluate.
// "False" needs only the atom's conservative world box (solid.cpp widens it
// far beyond float64 rounding); "true" is proven for untransformed boxes and
// bounds checks (exact comparisons) and spheres (farthest brick corner inside
// the threshold with a relative margin far above the rounding of the sum).
// The program is read through the given pointer types (LDS, global or
// constant memory on the device; plain pointers on the host).
template <class BoxP>
MC3D_HD inline bool box_misses(BoxP boxes, uint32_t i, const double *lo, const double *hi) {
    return hi[0] < boxes[i].lo[0] || lo[0] > boxes[i].hi[0] || hi[1] < boxes[i].lo[1] || lo[1] > boxes[i].hi[1] ||
           hi[2] < boxes[i].lo[2] || lo[2] > boxes[i].hi[2];
}

template <class F64P, class BoxP>
MC3D_HD inline int classify_atom(F64P f64, BoxP boxes, uint32_t pc, Insn in, const double *lo, const double *hi) {
    if (box_misses(boxes, pc, lo, hi)) return 0;
    if (insn_frame(in.x) == 0) {
        const auto a = f64 + in.y;
        const uint32_t kind = insn_kind(in.x);
        if (kind == A_RECT) {
            if (in_lo(lo[0], a[0]) && in_lo(lo[1], a[1]) && in_lo(lo[2], a[2]) && in_hi(hi[0], a[3]) &&
                in_hi(hi[1], a[4]) && in_hi(hi[2], a[5]))
                return 1;
        } else if (kind == A_SPHERE) {
            double s = 0;
            for (int i = 0; i < 3; i++) {
                double m = fmax(fabs(lo[i] - a[i]), fabs(hi[i] - a[i]));
                s += m * m;
            }
            if (s * (1 + 1e-12) <= a[4]) return 1;
        }
    }
    return 2;
}

constexpr uint32_t TAG_EXPR = 0u, TAG_FALSE = 1u << 30, TAG_TRUE = 2u << 30, TAG_MASK = 3u << 30;

// One instruction of the fold (S: the stack, sp its depth).
template <class BoxP, class F64P>
MC3D_HD inline void fold_step(BoxP boxes, F64P f64, uint32_t pc, Insn in, uint32_t *S, int &sp, const double *lo,
                              const double *hi, Insn *o, uint32_t cap, uint32_t &n, bool &over) {
    const uint32_t op = insn_op(in.x);
    if (op == OP_ATOM) {
        int c = classify_atom(f64, boxes, pc, in, lo, hi);
        if (c == 2) {
            if (n < cap) o[n] = in;
            else over = true;
            S[sp++] = TAG_EXPR | n;
            n++;
        } else {
            S[sp++] = c ? TAG_TRUE : TAG_FALSE;
        }
    } else if (op == OP_TRUE || op == OP_FALSE) {
        S[sp++] = op == OP_TRUE ? TAG_TRUE : TAG_FALSE;
    } else if (op == OP_NOT) {
        uint32_t a = S[sp - 1];
        if (a == TAG_TRUE) S[sp - 1] = TAG_FALSE;
        else if (a == TAG_FALSE) S[sp - 1] = TAG_TRUE;
        else {
            if (n < cap) o[n] = in;
            else over = true;
            n++;
        }
    } else {
        const uint32_t bb = S[--sp], aa = S[--sp];
        const bool aE = (aa & TAG_MASK) == TAG_EXPR, bE = (bb & TAG_MASK) == TAG_EXPR;
        const bool aT = aa == TAG_TRUE, aF = aa == TAG_FALSE, bT = bb == TAG_TRUE, bF = bb == TAG_FALSE;
        // start of the operands' code (constants emitted nothing)
        const uint32_t first = aE ? (aa & ~TAG_MASK) : (bE ? (bb & ~TAG_MASK) : n);
        uint32_t res;
        bool emit = false, emit_not = false;
        if (op == OP_OR) {
            if (aT || bT) res = TAG_TRUE, n = first;
            else if (aF) res = bb;
            else if (bF) res = aa;
            else res = aa, emit = true;
        } else if (op == OP_AND) {
            if (aF || bF) res = TAG_FALSE, n = first;
            else if (aT) res = bb;
            else if (bT) res = aa;
            else res = aa, emit = true;
        } else if (op == OP_ANDNOT) {  // a & !b
            if (aF || bT) res = TAG_FALSE, n = first;
            else if (bF) res = aa;
            else if (aT) res = bb, emit_not = true;  // b's code is last
            else res = aa, emit = true;
        } else {  // OP_RANDNOT: !a & b
            if (aT || bF) res = TAG_FALSE, n = first;
            else if (aF) res = bb;
            else if (bT) res = aa, emit_not = true;  // a's code is last
            else res = aa, emit = true;
        }
        if (emit || emit_not) {
            if (n < cap) o[n] = emit ? in : Insn{OP_NOT, 0};
            else over = true;
            n++;
        }
        S[sp++] = res;
    }
}

// Folds the whole program against one brick's box (lo, hi), instruction by
// instruction; the folded code goes to o[0, cap) (`over` if it needs more),
// the root's tag to S[0].
template <class CodeP, class BoxP, class F64P>
MC3D_HD inline void plan_fold(CodeP code, BoxP boxes, F64P f64, uint32_t ncode, uint32_t *S, const double *lo,
                              const double *hi, Insn *o, uint32_t cap, uint32_t &n, bool &over) {
    int sp = 0;
    for (uint32_t pc = 0; pc < ncode; ++pc)
        fold_step(boxes, f64, pc, Insn{code[pc].x, code[pc].y}, S, sp, lo, hi, o, cap, n, over);
}

// A value every lane of a wave holds (kept in a scalar register on the
// device; the host walks one brick at a time).
#if defined(__HIP_DEVICE_COMPILE__)
#define MC3D_UNIFORM(x) ((uint32_t)__builtin_amdgcn_readfirstlane((int)(x)))
#else
#define MC3D_UNIFORM(x) (x)
#endif

MC3D_HD inline bool node_misses(const PlanNode &nd, const double *lo, const double *hi) {
    return hi[0] < nd.lo[0] || lo[0] > nd.hi[0] || hi[1] < nd.lo[1] || lo[1] > nd.hi[1] || hi[2] < nd.lo[2] ||
           lo[2] > nd.hi[2];
}

// classify_atom from a node record (the same tests on the same values).
MC3D_HD inline int classify_node(const PlanNode &nd, const double *lo, const double *hi) {
    if (node_misses(nd, lo, hi)) return 0;
    if (insn_frame(nd.x) == 0) {
        const uint32_t kind = insn_kind(nd.x);
        const double *a = nd.prm;
        if (kind == A_RECT) {
            if (in_lo(lo[0], a[0]) && in_lo(lo[1], a[1]) && in_lo(lo[2], a[2]) && in_hi(hi[0], a[3]) &&
                in_hi(hi[1], a[4]) && in_hi(hi[2], a[5]))
                return 1;
        } else if (kind == A_SPHERE) {
            double s = 0;
            for (int i = 0; i < 3; i++) {
                double m = fmax(fabs(lo[i] - a[i]), fabs(hi[i] - a[i]));
                s += m * m;
            }
            if (s * (1 + 1e-12) <= a[3]) return 1;
        }
    }
    return 2;
}

// A node fetched whole (one 128-byte read: LDS broadcast, or scalar loads
// through a constant-address-space pointer), its instruction and link made
// wave-uniform.
template <class NodeP>
MC3D_HD inline PlanNode fetch_node(NodeP nodes, uint32_t i) {
    PlanNode nd;  // field by field: the pointer may be in any address space
    nd.x = MC3D_UNIFORM(nodes[i].x), nd.y = MC3D_UNIFORM(nodes[i].y);
    nd.link = MC3D_UNIFORM(nodes[i].link), nd.cfalse = MC3D_UNIFORM(nodes[i].cfalse);
    for (int k = 0; k < 3; k++) nd.lo[k] = nodes[i].lo[k], nd.hi[k] = nodes[i].hi[k];
    for (int k = 0; k < 6; k++) nd.prm[k] = nodes[i].prm[k];
    return nd;
}

// The fold with subtrees skipped, over node records (plan_nodes): at a leaf,
// the nodes whose subtrees start there are tried outermost first; the first
// one for which all_miss(node) holds is replaced by its all-false constant
// and the walk continues after it.  all_miss must hold only when the node's
// box misses this brick (on the device: every brick of the wave, so that
// the walk stays the same for all lanes).  Everything else is fold_step's
// rule on the node's instruction, box and params, so the folded code, the
// root and the overflow flag are plan_fold's.
template <class NodeP, class AllMiss>
MC3D_HD inline void plan_fold_nodes(NodeP nodes, uint32_t ncode, uint32_t *S, const double *lo, const double *hi,
                                    Insn *o, uint32_t cap, uint32_t &n, bool &over, AllMiss &&all_miss) {
    int sp = 0;
    uint32_t pc = 0;
    while (pc < ncode) {
        PlanNode cur = fetch_node(nodes, pc);
        const uint32_t op = insn_op(cur.x);
        if (op == OP_ATOM || op == OP_TRUE || op == OP_FALSE) {
            uint32_t nd = cur.link;
            bool skip = false;
            while (nd != pc) {
                const PlanNode t = fetch_node(nodes, nd);
                if (all_miss(t)) {  // subtree [pc, nd] is constant in these bricks
                    S[sp++] = t.cfalse ? TAG_TRUE : TAG_FALSE;
                    pc = nd + 1;
                    skip = true;
                    break;
                }
                nd = t.link;
            }
            if (skip) continue;
        }
        const Insn in{cur.x, cur.y};
        if (op == OP_ATOM) {
            const int c = classify_node(cur, lo, hi);
            if (c == 2) {
                if (n < cap) o[n] = in;
                else over = true;
                S[sp++] = TAG_EXPR | n;
                n++;
            } else {
                S[sp++] = c ? TAG_TRUE : TAG_FALSE;
            }
        } else {
            // not an atom: the operators' rules (no box involved)
            fold_step((const AtomBox *)nullptr, (const double *)nullptr, pc, in, S, sp, lo, hi, o, cap, n, over);
        }
        ++pc;
    }
}

}  // namespace mc3d
