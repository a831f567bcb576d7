/* ORACLE — test infrastructure only (see oracle.h).
 *
 * The fixed-order reduction of compiler/resident/reductions.py:1-45 (SLOTS = 65536, FINAL = 256,
 * DIRECT = 65536 at :66-72): n > DIRECT iterations are dealt to 65536 slots (slot s takes s, s+65536, ...
 * in order), the slots are folded into 256 values (f takes f, f+256, ...), and the 256 values are folded
 * in index order into the initial value; n <= DIRECT goes straight to 256 values. max/min fold as the
 * reference's ordered maximum/minimum (compiler/runtime/numeric.hpp:48-58).
 * Pinned by: max, min and max-magnitude equal the serial result (order-free operations); an integer-
 * valued sum below 2^53 equals the exact serial sum; and two hand-derived known answers of real sums whose
 * ties round differently in every other order (serial, contiguous slots or folds, reversed folds, the
 * initial value last) fix the slot, fold and host orders of both the direct and the two-level path
 * (tests/test_reduce.py ORDER_KATS).
 */
#include <math.h>
#include <stdlib.h>

enum { R_SUM = 0, R_SUMSQ = 1, R_MAX = 2, R_MIN = 3, R_MAXABS = 4 };
#define SLOTS 65536
#define FINAL 256

static double comb(int op, double acc, double x)
{
    switch (op) {
    case R_SUM: return acc + x;
    case R_SUMSQ: return acc + x * x;
    case R_MAX: return acc > x ? acc : x;
    case R_MIN: return acc < x ? acc : x;
    default: {
        double m = x < 0.0 ? -x : x;
        return acc > m ? acc : m;
    }
    }
}

static double ident(int op)
{
    return op == R_MAX ? -HUGE_VAL : (op == R_MIN ? HUGE_VAL : 0.0);
}

static int fold_op(int op) { return op == R_SUMSQ ? R_SUM : (op == R_MAXABS ? R_MAX : op); }

static double item(const double *x, long long t, long long n1, long long n2, long long ld1, long long ld2)
{
    long long i = t % n1, r = t / n1;
    return x[i + (r % n2) * ld1 + (r / n2) * ld2];
}

double orc_fixed_order_reduce(const double *x, long long n1, long long n2, long long n3, long long ld1,
                              long long ld2, int op, double initial)
{
    long long n = n1 * n2 * n3;
    double fin[FINAL];
    if (n <= SLOTS) {
        for (int f = 0; f < FINAL; ++f) {
            double acc = ident(op);
            for (long long t = f; t < n; t += FINAL) acc = comb(op, acc, item(x, t, n1, n2, ld1, ld2));
            fin[f] = acc;
        }
    } else {
        double *slots = (double *)malloc(sizeof(double) * SLOTS);
#pragma omp parallel for
        for (int s = 0; s < SLOTS; ++s) {
            double acc = ident(op);
            for (long long t = s; t < n; t += SLOTS) acc = comb(op, acc, item(x, t, n1, n2, ld1, ld2));
            slots[s] = acc;
        }
        int fop = fold_op(op);
        for (int f = 0; f < FINAL; ++f) {
            double acc = ident(fop);
            for (int t = f; t < SLOTS; t += FINAL) acc = comb(fop, acc, slots[t]);
            fin[f] = acc;
        }
        free(slots);
    }
    int hop = fold_op(op);
    double acc = initial;
    for (int f = 0; f < FINAL; ++f) acc = comb(hop, acc, fin[f]);
    return acc;
}
