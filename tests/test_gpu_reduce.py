"""Top-k per row and threshold compaction on the GPU (libcsamd.so) against
numpy orderings of the same matrices.  Bar: bit-exact (scores are copied,
not computed)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def ref_topk(score, k):
    """(score desc, column asc), NaN last (by column), padded with NaN / -1."""
    rows, cols = score.shape
    out_s = np.full((rows, k), np.nan, np.float32)
    out_i = np.full((rows, k), -1, np.int32)
    c = np.arange(cols)
    for r in range(rows):
        s = score[r]
        nan = np.isnan(s)
        order = np.lexsort((c, -np.where(nan, 0, s), nan))
        take = order[:k]
        out_s[r, :len(take)] = s[take]
        out_i[r, :len(take)] = take
    return out_s, out_i


@pytest.mark.parametrize("rows,cols", [(3, 100), (7, 4096), (5, 4097), (4, 20000), (2, 123457)])
@pytest.mark.parametrize("k", [1, 4, 10, 16])
def test_topk_rows_matches_numpy_order(gpu, rows, cols, k):
    import torch

    from cudams.device import topk_rows

    rng = np.random.default_rng(rows * 1000 + cols + k)
    # coarse values: many exact ties across chunks; some NaN
    s = (rng.integers(0, 50, size=(rows, cols)) / 49.0).astype(np.float32)
    s[rng.random((rows, cols)) < 0.01] = np.nan
    m = rng.integers(0, 100, size=(rows, cols)).astype(np.int32)
    ld = cols + 3
    S = torch.full((rows, ld), 5.0, dtype=torch.float32, device="cuda")
    M = torch.zeros((rows, ld), dtype=torch.int32, device="cuda")
    S[:, :cols] = torch.from_numpy(s).cuda()
    M[:, :cols] = torch.from_numpy(m).cuda()
    ts, ti, tm = topk_rows(S, k, M, rows=rows, cols=cols)
    torch.cuda.synchronize()
    ws, wi = ref_topk(s, k)
    ti = ti.cpu().numpy()
    np.testing.assert_array_equal(ti, wi)
    np.testing.assert_array_equal(ts.cpu().numpy(), ws)
    want_m = np.where(wi >= 0, m[np.arange(rows)[:, None], np.maximum(wi, 0)], 0)
    np.testing.assert_array_equal(tm.cpu().numpy(), want_m)


@pytest.mark.parametrize("rows,cols,k", [(64, 70000, 10), (240, 140000, 10), (240, 140000, 3)])
def test_topk_rows_threshold_edge_rows(gpu, rows, cols, k):
    """Blocks that rank several 4096-column batches each (the first pass
    inserts only values at or above a per-wave bound), on the rows that
    bound handles specially: mostly zeros with tied positives, all zeros (the
    bound is 0 and every value ties), one value everywhere, fewer than k real
    values among NaN, all NaN, and plain random rows."""
    import torch

    from cudams.device import topk_rows

    rng = np.random.default_rng(rows + cols + k)
    s = np.empty((rows, cols), np.float32)
    for r in range(rows):
        kind = r % 6
        if kind == 0:
            x = (rng.integers(1, 20, cols) / 19.0).astype(np.float32)
            x[rng.random(cols) < 0.88] = 0.0
        elif kind == 1:
            x = np.zeros(cols, np.float32)
        elif kind == 2:
            x = np.full(cols, 0.5, np.float32)
        elif kind == 3:
            x = np.full(cols, np.nan, np.float32)
            x[rng.choice(cols, min(k - 1, 2), replace=False)] = rng.random(min(k - 1, 2))
        elif kind == 4:
            x = np.full(cols, np.nan, np.float32)
        else:
            x = rng.random(cols).astype(np.float32)
        s[r] = x
    m = rng.integers(0, 100, size=(rows, cols)).astype(np.int32)
    S = torch.from_numpy(s).cuda()
    M = torch.from_numpy(m).cuda()
    print(f"top-k {rows}x{cols} k={k}", flush=True)
    ts, ti, tm = topk_rows(S, k, M)
    torch.cuda.synchronize()
    ws, wi = ref_topk(s, k)
    np.testing.assert_array_equal(ti.cpu().numpy(), wi)
    np.testing.assert_array_equal(ts.cpu().numpy(), ws)
    want_m = np.where(wi >= 0, m[np.arange(rows)[:, None], np.maximum(wi, 0)], 0)
    np.testing.assert_array_equal(tm.cpu().numpy(), want_m)


def test_topk_fewer_columns_than_k(gpu):
    import torch

    from cudams.device import topk_rows

    s = torch.tensor([[0.5, 0.7, 0.1]], dtype=torch.float32, device="cuda")
    ts, ti, _ = topk_rows(s, 5)
    assert ti.cpu().numpy().tolist() == [[1, 0, 2, -1, -1]]
    assert np.all(np.isnan(ts.cpu().numpy()[0, 3:]))


def test_threshold_compact_matches_numpy(gpu):
    import torch

    from cudams.device import threshold_compact

    rng = np.random.default_rng(4)
    s = rng.random((37, 1001)).astype(np.float32)
    m = rng.integers(0, 9, size=s.shape).astype(np.int32)
    o = (rng.random(s.shape) < 0.1).astype(np.uint8)
    S, M, O = (torch.from_numpy(x).cuda() for x in (s, m, o))
    r, c, sc, mm, oo = threshold_compact(S, M, O, 0.9, row0=100)
    rr, cc = np.nonzero(s >= np.float32(0.9))
    np.testing.assert_array_equal(r, rr + 100)
    np.testing.assert_array_equal(c, cc)
    np.testing.assert_array_equal(sc, s[rr, cc])
    np.testing.assert_array_equal(mm, m[rr, cc])
    np.testing.assert_array_equal(oo, o[rr, cc])


@pytest.mark.parametrize("mz_power,intensity_power", [(0.0, 1.0), (0.5, 1.0), (0.0, 0.5), (1.5, 2.0)])
def test_device_norms_equal_host_norms(gpu, mz_power, intensity_power):
    """cs_spectra_norms (DeviceSpectra's norms) equals the host's left-to-right
    float64 sum bit for bit, empty spectra (norm 0) included."""
    from cudams import synthetic
    from cudams.spectra import DeviceSpectra, PackedSpectra, power_factors

    p = synthetic.lognormal_spectra(3000, 17, cluster_fraction=0.2)
    mzs = [p.mz[p.offsets[i]:p.offsets[i + 1]] for i in range(p.n_spectra)]
    its = [p.intensities[p.offsets[i]:p.offsets[i + 1]] for i in range(p.n_spectra)]
    for k in (0, 7, 2999):  # empty spectra at the ends and inside
        mzs[k], its[k] = mzs[k][:0], its[k][:0]
    p = PackedSpectra.from_arrays(mzs, its)
    d = DeviceSpectra(p, mz_power, intensity_power, gpu)
    want = power_factors(p, mz_power, intensity_power)[2]
    got = d.norm.cpu().numpy()[:p.n_spectra]
    assert got[0] == 0.0 and got[7] == 0.0 and got[2999] == 0.0
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("bad", ["none", "unsorted", "offsets_past_end", "offsets_decrease",
                                 "max_peaks", "max_abs_mz", "nan", "exact_unsorted"])
def test_device_validator_flags_each_fault(gpu, bad):
    """cs_validate_spectra over a set whose device arrays were corrupted after
    packing: each kind of corruption raises its own fault bit, and only that
    one (the compute entry points take these properties on trust)."""
    import torch

    from cudams import _native
    from cudams.device import check_spectra_on_device, validate_spectra
    from cudams.spectra import DeviceSpectra, PackedSpectra

    mzs = [[100.0, 101.0, 102.0], [50.0, 60.0], [], [70.0, 70.5, 71.0, 80.0]]
    if bad == "exact_unsorted":
        mzs[0] = [100.0, 100.1, 102.0]  # not float32 numbers: mz_exact is kept
    p = PackedSpectra.from_arrays(mzs, [[1.0] * len(m) for m in mzs])
    S = DeviceSpectra(p, 0.0, 1.0, "cuda")
    want = {"none": 0, "unsorted": _native.CS_FAULT_MZ_ORDER,
            "offsets_past_end": _native.CS_FAULT_OFFSETS,
            # [0, 3, 5, 4, 9]: spectrum 2 runs backwards, spectrum 3 spans 5 peaks
            "offsets_decrease": _native.CS_FAULT_OFFSETS | _native.CS_FAULT_LENGTH,
            "max_peaks": _native.CS_FAULT_LENGTH,
            "max_abs_mz": _native.CS_FAULT_MZ_RANGE, "nan": _native.CS_FAULT_MZ_RANGE,
            "exact_unsorted": _native.CS_FAULT_MZ_ORDER}[bad]
    if bad == "unsorted":
        S.mz[5] = 75.0
    elif bad == "offsets_past_end":
        S.offsets[4] = 10 ** 12
    elif bad == "offsets_decrease":
        S.offsets[3] = 4
    elif bad == "max_peaks":
        S.max_peaks = 3
    elif bad == "max_abs_mz":
        S.max_abs_mz = 90.0
    elif bad == "nan":
        S.mz[3] = float("nan")
    elif bad == "exact_unsorted":
        assert S.mz_exact is not None
        S.mz_exact[1] = 99.0
    torch.cuda.synchronize()
    print(f"validator case {bad}", flush=True)
    assert validate_spectra(S) == want
    if want:
        with pytest.raises(ValueError, match="corrupt spectrum set"):
            check_spectra_on_device(S)
    else:
        check_spectra_on_device(S)


@pytest.mark.parametrize("rows,cols,pad", [(1, 1, 0), (7, 5, 3), (33, 127, 1), (256, 300, 0)])
def test_pack_records_matches_numpy_structured_array(gpu, rows, cols, pad):
    """cs_pack_records gives byte for byte the memory numpy builds for the
    structured dtype from the three planes (padded leading dimension, record
    counts that are not a multiple of four)."""
    import torch

    from cudams.device import pack_records

    rng = np.random.default_rng(rows * 1000 + cols)
    ld = cols + pad
    s = rng.standard_normal((rows, ld)).astype(np.float32)
    m = rng.integers(-2 ** 31, 2 ** 31 - 1, (rows, ld)).astype(np.int32)
    o = rng.integers(0, 256, (rows, ld)).astype(np.uint8)
    dev = [torch.from_numpy(x).cuda() for x in (s, m, o)]
    out = torch.full((9 * rows * cols + 3,), 0xAB, dtype=torch.uint8, device="cuda")
    print(f"pack_records {rows}x{cols} ld {ld}", flush=True)
    pack_records(dev[0][:, :cols], dev[1][:, :cols], dev[2][:, :cols], out, rows, cols)
    got = out.cpu().numpy()
    want = np.rec.fromarrays([s[:, :cols], m[:, :cols], o[:, :cols]],
                             dtype=[("score", "<f4"), ("matches", "<i4"), ("overflow", "u1")])
    assert got[:9 * rows * cols].tobytes() == want.tobytes()
    assert (got[9 * rows * cols:] == 0xAB).all()
