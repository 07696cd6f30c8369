"""Output files of a render, through the C ABI on the host (no GPU): the
OpenEXR film (the reference writes HALF RGB through OpenImageIO,
src/image.cpp:45-58) and the statistics text (Stats::output,
src/integrators/delayed/delayed.cpp:311-324).

The EXR files are read back by tests/exr_read.py, a reader written from the
published file layout; the float -> half conversion is checked bit for bit
against numpy's IEEE binary16 rounding (round to nearest even) over ties,
subnormals, overflow and the special values."""
import numpy as np
import pytest

from drmlt import capi
from exr_read import read_exr


def _halves(vals, tmp_path):
    """The library's binary16 of each float, through an uncompressed one-row
    EXR (the R channel)."""
    vals = np.asarray(vals, np.float32)
    img = np.zeros((1, len(vals), 3), np.float32)
    img[0, :, 0] = vals
    p = tmp_path / "row.exr"
    capi.write_exr(p, img, capi.EXR_NONE)
    ch, _ = read_exr(p)
    return ch["R"][0]


def test_float_to_half_is_ieee_round_to_nearest_even(tmp_path):
    rng = np.random.default_rng(5)
    mags = np.exp2(rng.uniform(-30, 17, 20000)).astype(np.float32)
    vals = np.concatenate([mags, -mags[:5000]])
    # exact ties and near-ties around the half grid, normal and subnormal
    ulp_n = np.float32(2.0 ** -10)
    base = np.float32(1.0) + np.arange(64, dtype=np.float32) * ulp_n
    ties = np.concatenate([base + ulp_n / 2, base + ulp_n / 2 + np.float32(2 ** -23), base + ulp_n / 2 - np.float32(2 ** -23)])
    sub = np.arange(1, 2000, dtype=np.float32) * np.float32(2.0 ** -25)  # multiples of half a subnormal step
    special = np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 65536.0, 1e30, -1e30, np.inf, -np.inf, 2.0 ** -14,
                        2.0 ** -14 - 2.0 ** -25, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26, 2.0 ** -26, 1e-30], np.float32)
    vals = np.concatenate([vals, ties, sub, special]).astype(np.float32)
    got = _halves(vals, tmp_path)
    with np.errstate(over="ignore"):
        want = vals.astype(np.float16)
    assert np.array_equal(got.view(np.uint16), want.view(np.uint16))
    nan = _halves(np.array([np.nan, -np.nan], np.float32), tmp_path)
    assert np.isnan(nan).all()


@pytest.mark.parametrize("comp", [capi.EXR_NONE, capi.EXR_ZIP])
def test_exr_layout_and_round_trip(tmp_path, comp):
    """A 37 x 21 film (ZIP: one full and one partial 16-line chunk): the
    header carries the required attributes with the reference's layout
    (B, G, R as HALF, increasing y, the data window of the image), and the
    pixels read back as the binary16 of the written floats."""
    rng = np.random.default_rng(comp)
    img = (rng.lognormal(-2, 2, (21, 37, 3)) * (rng.uniform(0, 1, (21, 37, 3)) > 0.2)).astype(np.float32)
    img[3, 4] = [0.0, 70000.0, 1e-6]
    p = tmp_path / f"film{comp}.exr"
    capi.write_exr(p, img, comp)
    ch, attrs = read_exr(p)
    req = {"channels": "chlist", "compression": "compression", "dataWindow": "box2i", "displayWindow": "box2i",
           "lineOrder": "lineOrder", "pixelAspectRatio": "float", "screenWindowCenter": "v2f", "screenWindowWidth": "float"}
    assert {k: attrs[k][0] for k in req} == req
    assert attrs["compression"][1][0] == comp and attrs["lineOrder"][1][0] == 0
    assert list(ch) == ["B", "G", "R"] and all(v.dtype == np.float16 for v in ch.values())
    with np.errstate(over="ignore"):
        want = img.astype(np.float16)
    for k, c in enumerate("RGB"):
        assert np.array_equal(ch[c].view(np.uint16), want[..., k].view(np.uint16)), c


def test_zip_smaller_than_raw_on_smooth_images(tmp_path):
    y, x = np.mgrid[0:64, 0:96].astype(np.float32)
    img = np.stack([x / 96, y / 64, (x + y) / 160], -1).astype(np.float32)
    capi.write_exr(tmp_path / "z.exr", img, capi.EXR_ZIP)
    capi.write_exr(tmp_path / "n.exr", img, capi.EXR_NONE)
    assert (tmp_path / "z.exr").stat().st_size < 0.6 * (tmp_path / "n.exr").stat().st_size
    a, _ = read_exr(tmp_path / "z.exr")
    b, _ = read_exr(tmp_path / "n.exr")
    assert all(np.array_equal(a[c], b[c]) for c in "BGR")


def test_exr_argument_errors(tmp_path):
    with pytest.raises(capi.DrmltError):
        capi.write_exr(tmp_path / "x.exr", np.zeros((2, 2, 3), np.float32), 7)  # unsupported compression
    with pytest.raises(capi.DrmltError):
        capi.write_exr(tmp_path / "no_such_dir" / "x.exr", np.zeros((2, 2, 3), np.float32))
    with pytest.raises(ValueError):
        capi.write_exr(tmp_path / "x.exr", np.zeros((2, 2), np.float32))


def _reference_text(s, swapped):
    """delayed.cpp:311-324 for the counts s, with std::ostream's default float
    formatting (precision 6, the shortest of fixed and scientific)."""
    f = lambda a, b: np.float32(a) / np.float32(b) if b else (np.float32("nan") if a == 0 else np.float32("inf"))

    def g(v):
        if np.isnan(v):
            return "-nan"  # 0/0 in float on x86-64 is the negative default NaN, which libstdc++ prints so
        return f"{float(v):g}"

    iso = (g(f(s["stage2_iso_accepted"], s["stage2_iso"])), s["stage2_iso"])
    h2 = (g(f(s["stage2_h2mc_accepted"], s["stage2_h2mc"])), s["stage2_h2mc"])
    first, second = (h2, iso) if swapped else (iso, h2)
    return ("\n"
            "Rendering Statistics \n"
            f"    Large step mutations acceptance ratio                        : {g(f(s['large_step_accepted'], s['large_step']))} ({s['large_step']})\n"
            f"    First stage mutations acceptance ratio                       : {g(f(s['stage1_accepted'], s['stage1']))} ({s['stage1']})\n"
            f"    Second stage mutations acceptance ratio, isotropic component : {first[0]} ({first[1]})\n"
            f"    Second stage mutations acceptance ratio, H2MC component      : {second[0]} ({second[1]})\n"
            "\n")


@pytest.mark.parametrize("swapped", [True, False])
def test_stats_text_is_the_references(swapped):
    cases = [
        dict(large_step=274973, large_step_accepted=2054, stage1=642531, stage1_accepted=239074, stage2_iso=201593,
             stage2_iso_accepted=99267, stage2_h2mc=201864, stage2_h2mc_accepted=87191),
        dict(large_step=3, large_step_accepted=1, stage1=7, stage1_accepted=7, stage2_iso=0, stage2_iso_accepted=0,
             stage2_h2mc=9, stage2_h2mc_accepted=0),
        dict(large_step=10 ** 12, large_step_accepted=1, stage1=1, stage1_accepted=0, stage2_iso=6, stage2_iso_accepted=5,
             stage2_h2mc=2 ** 40, stage2_h2mc_accepted=2 ** 39),
    ]
    for s in cases:
        assert capi.stats_text(s, reference_labels=swapped) == _reference_text(s, swapped)
    # the reference's swap: its "isotropic" line carries the H2MC count
    t = capi.stats_text(cases[0], reference_labels=True).splitlines()
    assert t[4].endswith("(201864)") and "isotropic" in t[4]


def test_mcmc_stats_text():
    """The mcmc integrator's Stats::output (mcmc.cpp:76-86): large and small
    steps; lens steps are never made by its SmallStep (mutation.cpp:126-165),
    so their line reads 0 of 0."""
    s = dict(large_step=300, large_step_accepted=20, stage1=700, stage1_accepted=350, stage2_iso=0, stage2_iso_accepted=0,
             stage2_h2mc=0, stage2_h2mc_accepted=0)
    want = ("\n"
            "Rendering Statistics \n"
            f"    Large step mutations acceptance ratio       : {float(np.float32(20) / np.float32(300)):g} (300)\n"
            "    Small step mutations acceptance ratio       : 0.5 (700)\n"
            "    Lens mutations acceptance ratio    :                    -nan (0)\n"
            "\n")
    assert capi.stats_text(s, integrator="mcmc") == want


def test_h2mc_stats_text():
    """h2mc's Stats::output (h2mc.cpp:140-150) has the mcmc layout; its small
    steps are the isotropic mixing moves plus the H2MC moves."""
    s = dict(large_step=10, large_step_accepted=1, stage1=20, stage1_accepted=10, stage2_iso=0, stage2_iso_accepted=0,
             stage2_h2mc=80, stage2_h2mc_accepted=30)
    t = capi.stats_text(s, integrator="h2mc").splitlines()
    assert t[3] == "    Small step mutations acceptance ratio       : 0.4 (100)"
    # lens mutations on their own line
    t = capi.stats_text(dict(s, lens_step=40, lens_step_accepted=30), integrator="h2mc").splitlines()
    assert t[3] == "    Small step mutations acceptance ratio       : 0.4 (100)"
    assert t[4] == "    Lens mutations acceptance ratio    :                    0.75 (40)"
