// Fixed-order reductions shared by the reduction entry and the pressure solve's norms.
#pragma once
#include <hip/hip_runtime.h>

namespace ftc {

enum ReduceOp : int { kReduceSum = 0, kReduceSumSquares = 1, kReduceMax = 2, kReduceMin = 3, kReduceMaxAbs = 4 };
constexpr int kReduceSlots = 65536;  // compiler/resident/reductions.py:66 (SLOTS)
constexpr int kReduceFinal = 256;    // compiler/resident/reductions.py:67 (FINAL)
constexpr long long kReduceDirect = 65536;  // compiler/resident/reductions.py:69 (DIRECT)

// A column-major box of n1 x n2 x n3 doubles starting at x, with strides ld1 (j) and ld2 (k).
struct ReduceBox {
    const double *x;
    long long n1, n2, n3, ld1, ld2;
};

// Enqueues the slot and fold kernels; final256 receives the 256 values the host folds.
int reduce_launch(int op, const ReduceBox &b, double *slots, double *final256, hipStream_t stream);
// As reduce_launch, then one thread folds the 256 values in index order into `initial` and stores the
// result in device memory.
int reduce_launch_to_device(int op, const ReduceBox &b, double *slots, double *final256, double initial,
                            double *result, hipStream_t stream);
// Folds the 256 values in index order into `initial`.
double reduce_host_fold(int op, double initial, const double *final256);

}  // namespace ftc
