"""The projected LES step and the batched ensemble.

CPU: the oracle's projection leaves a divergence-free field (div grad is exactly the pressure solve's
7-point Laplacian), which pins the divergence and projection restatements.

GPU: the per-member FP64 ensemble against the oracle (LES step bit-exact, pressure within the direct
solve's rounding, so fields within 1e-12 of their scale); the batched FP64 ensemble bit for bit equal to
the per-member one (same kernels, same order); the mixed batched ensemble within 1e-10 of the FP64 one
(the north-star field error bound); the autotuner's cache round trip.
"""
import numpy as np
import pytest

from conftest import bits_equal, describe_mismatch

SIZES = [(40, 36, 12), (33, 47, 9), (64, 32, 16), (8, 6, 4)]
D, DT, CS = (10.0, 10.0, 10.0), 0.2, 0.17


def member_fields(fields, n, seed, d=D):
    return fields.les_fields(*n, *d, seed=seed)


def test_projected_oracle_step_is_divergence_free(oracle, fields):
    U, V, W, S = member_fields(fields, (24, 20, 10), 11)
    Un, Vn, Wn, Sn, phi = oracle.projected_step(U, V, W, S, *D, DT, CS)
    before = oracle.divergence(*oracle.les_step(U, V, W, S, *D, DT, CS)[:3], *D, DT)
    after = oracle.divergence(Un, Vn, Wn, *D, DT)
    assert np.abs(after).max() <= 1e-10 * np.abs(before).max()
    assert abs(phi.mean()) <= 1e-12 * np.abs(phi).max()


def run(hip, ens_mod, fields, mode, steps, tune=None, solves=None, sizes=SIZES, seed=100, projection=True,
        policy=None, d=D):
    e = ens_mod.Ensemble(hip, sizes, d, DT, CS, mode)
    try:
        if solves:
            e.solves(solves)
        if policy is not None:
            e.les_policy(policy)
        if not projection:
            e.projection(False)
        if tune is not None:
            e.autotune(tune)
        for m, n in enumerate(sizes):
            e.upload(m, *member_fields(fields, n, seed + m, d))
        e.step(steps)
        e.synchronize()
        out = [e.download(m) + ([e.pressure(m)] if projection else []) for m in range(len(sizes))]
        info = e.info()
    finally:
        e.destroy()
    return out, info


@pytest.fixture(scope="module")
def ens_mod(ftc):
    import __graft_entry__ as graft

    return graft._module("ensemble")


@pytest.mark.gpu
@pytest.mark.parametrize("steps", [1, 2])
def test_fp64_members_match_the_oracle(hip, ens_mod, oracle, fields, steps):
    got, info = run(hip, ens_mod, fields, ens_mod.FP64_PER_MEMBER, steps)
    assert info.launches_per_step == 8 * len(SIZES)
    for m, n in enumerate(SIZES):
        state = list(member_fields(fields, n, 100 + m))
        for _ in range(steps):
            *state, phi = oracle.projected_step(*state, *D, DT, CS)
        for name, g, w in zip("UVWS", got[m][:4], state):
            scale = np.abs(w).max()
            assert np.abs(g - w).max() <= 1e-12 * scale, (m, name, np.abs(g - w).max() / scale)
        assert np.abs(got[m][4] - phi).max() <= 1e-11 * np.abs(phi).max()


@pytest.mark.gpu
def test_batched_fp64_equals_per_member_fp64_bit_for_bit(hip, ens_mod, fields):
    ref, _ = run(hip, ens_mod, fields, ens_mod.FP64_PER_MEMBER, 3)
    got, info = run(hip, ens_mod, fields, ens_mod.FP64_BATCHED, 3)
    assert info.launches_per_step == 8
    for m in range(len(SIZES)):
        for g, w in zip(got[m], ref[m]):
            assert bits_equal(g, w)


# The FP64 transforms run as axis transforms (one LDS block of 64 columns or rows per workgroup) when every
# member's extent along the axis is at most 128, else as tiled contractions; both keep the operand roles and
# the order of the MFMAs over K, so they give the same bits. The second size list puts one member above 128
# in x (x transforms tiled, y transforms along the axis), the third one above 128 in y.
@pytest.mark.gpu
@pytest.mark.parametrize("sizes", [SIZES, [(150, 40, 8), (40, 36, 12), (33, 100, 6)],
                                   [(40, 140, 7), (64, 32, 10), (128, 17, 5)]])
def test_axis_transforms_equal_the_tiled_contractions_bit_for_bit(hip, ens_mod, fields, sizes, monkeypatch):
    print(f"ensemble transforms, sizes {sizes}", flush=True)
    ref, _ = run(hip, ens_mod, fields, ens_mod.FP64_BATCHED, 2, sizes=sizes)
    monkeypatch.setenv("FTC_ENS_TILED_TRANSFORMS", "1")
    tiled, _ = run(hip, ens_mod, fields, ens_mod.FP64_BATCHED, 2, sizes=sizes)
    for m in range(len(sizes)):
        for name, g, w in zip("UVWSP", ref[m], tiled[m]):
            assert bits_equal(g, w), (m, name, describe_mismatch(g, w))


@pytest.mark.gpu
def test_members_wider_than_an_axis_block_match_the_oracle(hip, ens_mod, oracle, fields):
    sizes = [(150, 40, 8), (40, 36, 12), (33, 100, 6)]
    got, _ = run(hip, ens_mod, fields, ens_mod.FP64_BATCHED, 1, sizes=sizes)
    for m, n in enumerate(sizes):
        *state, phi = oracle.projected_step(*member_fields(fields, n, 100 + m), *D, DT, CS)
        for name, g, w in zip("UVWS", got[m][:4], state):
            scale = np.abs(w).max()
            assert np.abs(g - w).max() <= 1e-12 * scale, (m, name, np.abs(g - w).max() / scale)
        assert np.abs(got[m][4] - phi).max() <= 1e-11 * np.abs(phi).max()


@pytest.mark.gpu
def test_mixed_batched_is_within_1e10_of_fp64(hip, ens_mod, fields):
    ref, _ = run(hip, ens_mod, fields, ens_mod.FP64_PER_MEMBER, 2)
    got, info = run(hip, ens_mod, fields, ens_mod.MIXED_BATCHED, 2)
    assert info.solves == 3
    for m in range(len(SIZES)):
        for name, g, w in zip("UVWSP", got[m], ref[m]):
            err = np.abs(g - w).max() / np.abs(w).max()
            assert err <= 1e-10, (m, name, err)


@pytest.mark.gpu
def test_autotuned_ensemble_reuses_its_cache_and_keeps_fp64_results(hip, ens_mod, fields, tmp_path):
    cache = str(tmp_path / "tune.tsv")
    ref, _ = run(hip, ens_mod, fields, ens_mod.FP64_BATCHED, 2)
    got, info = run(hip, ens_mod, fields, ens_mod.FP64_BATCHED, 2, tune=cache)
    text = open(cache).read()
    assert "mode=1" in text
    again, info2 = run(hip, ens_mod, fields, ens_mod.FP64_BATCHED, 2, tune=cache)
    assert (info2.les_kchunk, info2.gemm_tile, info2.blocks_per_cu, info2.les_policy) == (
        info.les_kchunk, info.gemm_tile, info.blocks_per_cu, info.les_policy)
    assert "widths=" in text
    for m in range(len(SIZES)):
        for g, w, a in zip(got[m], ref[m], again[m]):
            assert bits_equal(g, w) and bits_equal(a, w)


# The seeded 64-member ensemble of the batched workload: nx, ny uniform in [32, 128], nz in [32, 96], so
# every member spans several k chunks of the LES kernel. Both batched modes are compared directly with the
# oracle after two steps: the LES stage alone bit for bit (it is FP64 in both modes), and the projected
# step within the FP64 direct solve's rounding (1e-12 of each velocity and scalar field's scale, 1e-11 for
# the pressure, as for the per-member runs above) or the mixed solve's bound (1e-10).
SEEDED_STEPS = 2


@pytest.fixture(scope="module")
def seeded64(oracle, fields):
    sizes = fields.ensemble_sizes(64)
    les, projected = [], []
    for m, n in enumerate(sizes):
        start = member_fields(fields, n, 500 + m)
        state = list(start)
        for _ in range(SEEDED_STEPS):
            state = list(oracle.les_step(*state, *D, DT, CS))
        les.append(state)
        state = list(start)
        for _ in range(SEEDED_STEPS):
            *state, phi = oracle.projected_step(*state, *D, DT, CS)
        projected.append(state + [phi])
    return sizes, les, projected


@pytest.mark.gpu
@pytest.mark.parametrize("mode,policy", [("FP64_BATCHED", 0), ("MIXED_BATCHED", 0), ("FP64_BATCHED", 1),
                                         ("FP64_BATCHED", 2), ("FP64_PER_MEMBER", 2)])
def test_seeded_64_members_les_stage_is_bit_exact_against_the_oracle(hip, ens_mod, fields, seeded64, mode, policy):
    # Tile policies: 0 every member in 64 x 4 tiles, 1 every member in 32 x 8 tiles, 2 each member in the
    # width that leaves fewer idle lanes (both widths in one step, one persistent launch per width).
    sizes, les, _ = seeded64
    print(f"seeded ensemble LES stage, {mode}, tile policy {policy}", flush=True)
    got, info = run(hip, ens_mod, fields, getattr(ens_mod, mode), SEEDED_STEPS, sizes=sizes, seed=500,
                    projection=False, policy=policy)
    assert info.les_policy == policy
    if mode != "FP64_PER_MEMBER":
        assert info.launches_per_step == (2 if policy == 2 else 1)
    for m, n in enumerate(sizes):
        for name, g, w in zip("UVWS", got[m], les[m]):
            assert bits_equal(g, w), (m, n, name, describe_mismatch(g, w))


@pytest.mark.gpu
@pytest.mark.parametrize("mode,bound,phi_bound", [("FP64_BATCHED", 1e-12, 1e-11), ("MIXED_BATCHED", 1e-10, 1e-10)])
def test_seeded_64_members_projected_step_matches_the_oracle(hip, ens_mod, fields, seeded64, mode, bound, phi_bound):
    sizes, _, projected = seeded64
    print(f"seeded ensemble projected step, {mode}", flush=True)
    got, info = run(hip, ens_mod, fields, getattr(ens_mod, mode), SEEDED_STEPS, sizes=sizes, seed=500)
    worst = 0.0
    for m, n in enumerate(sizes):
        for name, g, w in zip("UVWSP", got[m], projected[m]):
            err = np.abs(g - w).max() / np.abs(w).max()
            worst = max(worst, err)
            assert err <= (phi_bound if name == "P" else bound), (m, n, name, err)
    print(f"worst field error {worst:.2e}")


# Unequal spacings take the LES kernels that hold a coefficient per direction; members of 26 to 40 planes span
# two k chunks of the default length (24 planes), and both tile widths are used.
@pytest.mark.gpu
@pytest.mark.parametrize("policy", [0, 1])
def test_les_stage_with_unequal_spacings_is_bit_exact_against_the_oracle(hip, ens_mod, oracle, fields, policy):
    sizes, d = [(40, 36, 30), (33, 47, 26), (70, 20, 40), (8, 6, 5)], (10.0, 8.0, 6.0)
    print(f"ensemble LES stage, spacings {d}, tile policy {policy}", flush=True)
    got, info = run(hip, ens_mod, fields, ens_mod.FP64_BATCHED, 2, sizes=sizes, seed=700, projection=False,
                    policy=policy, d=d)
    assert info.les_policy == policy
    for m, n in enumerate(sizes):
        state = list(member_fields(fields, n, 700 + m, d))
        for _ in range(2):
            state = list(oracle.les_step(*state, *d, DT, CS))
        for name, g, w in zip("UVWS", got[m], state):
            assert bits_equal(g, w), (m, n, name, describe_mismatch(g, w))
