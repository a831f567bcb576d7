/* ORACLE — test infrastructure only.
 *
 * A plain-C restatement, on the CPU, of the loop nests on the resident LES path. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only as the checker or the timed
 * CPU baseline; the product path never calls it.
 *
 * Conventions shared by every routine:
 *   - arrays are column-major with padded extents (e1, e2, e3); indices are the Fortran 1-based ones of the
 *     reference loop nests, so interior cells are 2..n+1 with one ghost layer on each side;
 *   - the file is compiled with -ffp-contract=off -fno-fast-math: every + - * / sqrt is one IEEE-754
 *     binary64 operation, evaluated in the source order of the nest it restates (Fortran evaluates
 *     a+b+c as (a+b)+c and a*b*c as (a*b)*c).
 *
 * Parity status of each function is recorded at its definition ("pinned by ..." or "parity unpinned").
 */
#ifndef FTC_ORACLE_H
#define FTC_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#define ORC_IX(i, j, k) \
    ((size_t)((i) - 1) + (size_t)e1 * ((size_t)((j) - 1) + (size_t)e2 * (size_t)((k) - 1)))

#endif
