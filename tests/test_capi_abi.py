"""The C-ABI library loads without a GPU and exports every entry point the
public header declares; the header's struct layouts match the ctypes
mirror (checked by compiling a probe against include/drmlt_amd.h with gcc).
No compute call is made here."""
import ctypes as C
import os
import re
import subprocess

import pytest

from drmlt import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "drmlt_amd.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(drmlt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(capi.EXPORTS) == names


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "drmlt_amd.h"\n'
                   "int main(void){printf(\"%zu %zu %zu %zu %zu %zu\\n\", sizeof(drmlt_options), sizeof(drmlt_scene_desc),"
                   " sizeof(drmlt_path_rec), sizeof(drmlt_stats), offsetof(drmlt_options, seed), offsetof(drmlt_path_rec, lrr));"
                   "printf(\"%zu\\n\", sizeof(drmlt_schedule));"
                   "return 0;}\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == C.sizeof(capi.Options)
    assert out[1] == C.sizeof(capi.SceneDesc)
    assert out[2] == capi.PATH_REC.itemsize == 536
    assert out[3] == C.sizeof(capi.Stats)
    assert out[4] == capi.Options.seed.offset
    assert out[5] == capi.PATH_REC.fields["lrr"][1]
    assert out[6] == C.sizeof(capi.Schedule)


def test_defaults_are_the_references():
    o = capi.default_options()
    # src/dptoptions.h:10-36 and src/main.cpp:76
    assert (o.spp, o.num_init_samples, o.direct_spp, o.num_chains) == (16, 100000, 16, 1024)
    assert o.large_step_prob == pytest.approx(0.3) and o.perturb_std_dev == pytest.approx(0.01)
    assert o.aniso_perturb_prob == pytest.approx(0.5) and o.aniso_perturb_std_dev == pytest.approx(0.0025)
    assert o.roughness_threshold == pytest.approx(0.03) and o.discrete_std_dev == pytest.approx(0.01)
    assert o.type == 0 and o.max_derivatives_depth == 9 and o.use_light_coordinate_sampling == 1


def test_dpt_keys_map_to_options():
    o = capi.options_from_dpt({"type": "mira", "anisoperturbprob": 0.25, "maxdepth": 7, "numchains": 4096,
                               "acceptanceMap": True, "spp": 8})
    assert o.type == 1 and o.aniso_perturb_prob == pytest.approx(0.25) and o.max_depth == 7
    assert o.num_chains == 4096 and o.acceptance_map == 1 and o.spp == 8


def test_argument_errors_without_gpu():
    """Argument validation runs before any device call."""
    lib = capi.lib()
    assert lib.drmlt_ctx_create(None, None, None) == -1
    assert b"null" in lib.drmlt_last_error()
    assert lib.drmlt_options_default(None) == -1
