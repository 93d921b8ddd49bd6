"""The C-ABI library loads without a GPU, exports every entry point that
include/pcs_abi.h declares, and its host-only parts behave; the Python
declarations match the header."""
import ctypes as C
import pathlib
import re

import numpy as np

import oracle_lib as O
from part_conc_spt_abc_amd import _abi

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared():
    txt = (ROOT / "include" / "pcs_abi.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pcs_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(_abi.library_path()))
    names = declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_abi.PROTOTYPES), set(names) ^ set(_abi.PROTOTYPES)


def test_tracing_build_exports_the_same_entry_points():
    """lib/trace/libpcs.so (make trace) is a drop-in for the default library."""
    path = ROOT / "part_conc_spt_abc_amd" / "lib" / "trace" / "libpcs.so"
    if not path.exists():
        import pytest
        pytest.skip("the tracing build was not made")
    lib = C.CDLL(str(path))
    for n in declared():
        assert hasattr(lib, n), n
    assert lib.pcs_abi_version() == 1


def test_struct_layouts_match_header(tmp_path):
    """sizeof/offsetof of every ABI struct from the C compiler vs ctypes."""
    import subprocess

    structs = {n: getattr(_abi, n) for n in ("pcs_system", "pcs_proposal", "pcs_tuning", "pcs_server_config",
                                             "pcs_request", "pcs_completion", "pcs_prior")}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pcs_abi.h"', "int main(void){"]
    for n, cls in structs.items():
        lines.append(f'printf("{n} %zu\\n", sizeof({n}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{n}.{f} %zu\\n", offsetof({n}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), "-o", str(exe), str(src)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.splitlines())
    for n, cls in structs.items():
        assert int(out[n]) == C.sizeof(cls), n
        for f, _ in cls._fields_:
            assert int(out[f"{n}.{f}"]) == getattr(cls, f).offset, f"{n}.{f}"


def test_host_only_calls_without_gpu():
    lib = _abi.load()
    assert lib.pcs_abi_version() == 1
    assert lib.pcs_strerror(_abi.PCS_EINVAL) == b"invalid argument"
    # invalid system is rejected before any device work
    from part_conc_spt_abc_amd.model import Model

    m = Model(Lx=10, Ly=10, Lz=10, ax=5, ay=5, deltat=0.05, kmin=1, n_k=10, n_de=10, ub_de=1.0, n_species=2)
    h = C.c_void_p()
    assert lib.pcs_open(C.byref(m.to_abi()), None, 0, C.byref(h)) == _abi.PCS_EINVAL
    assert b"kmin" in lib.pcs_last_error()
    # NULL handles of the proposal step, the resident population, the
    # allocation-free entry point and the server are refused before any HIP call
    pr = _abi.pcs_prior()
    one = (C.c_double * 4)()
    out = (C.c_double * 4)()
    assert lib.pcs_batch_propose(None, C.byref(pr), one, one, one, 1, 1, None, 1, 0, 0, out) == _abi.PCS_EINVAL
    assert lib.pcs_batch_set_population(None, one, one, 1, 1) == _abi.PCS_EINVAL
    assert lib.pcs_batch_pmc_weights(None, one, out) == _abi.PCS_EINVAL
    assert lib.pcs_batch_read_proposals(None, None, None) == _abi.PCS_EINVAL
    assert lib.pcs_simulate_static(None, None, 0, None, None, 0, 0, None, None, None) == _abi.PCS_EINVAL
    us = C.c_double()
    assert lib.pcs_server_now_us(None, C.byref(us)) == _abi.PCS_EINVAL
    assert lib.pcs_server_set_record_capacity_factor(None, 1.0) == _abi.PCS_EINVAL
    n64 = C.c_uint64()
    assert lib.pcs_server_group_walked(None, C.byref(n64)) == _abi.PCS_EINVAL
    assert lib.pcs_server_overflow_reruns(None, C.byref(n64)) == _abi.PCS_EINVAL
    assert lib.pcs_batch_pmc_weights_async(None, one) == _abi.PCS_EINVAL
    assert lib.pcs_batch_pmc_weights_wait(None, out) == _abi.PCS_EINVAL
    # unknown server flags are refused before any device work
    cfg = _abi.pcs_server_config(0, 0, 0, 0x80, 0)
    assert lib.pcs_server_create(0, C.byref(cfg), C.byref(h)) == _abi.PCS_EINVAL
    assert b"flags" in lib.pcs_last_error()


def test_reference_hist_matches_oracle():
    rng = np.random.default_rng(3)
    K = rng.integers(1, 40, 5000).astype(np.uint64)
    DE = rng.exponential(2.0, 5000)
    DE[:10] = [0.0, 4.999999, 5.0, 7.0, 0.04, 0.05, 1e-300, 4.95, 2.5, 3.0]
    from part_conc_spt_abc_amd.gpu import reference_hist

    H, dropped = reference_hist(K, DE, 5.0, 30, 100)
    Ho, n, dro = O.reference_hist(K, DE, 5.0, 30, 100)
    assert np.array_equal(H, Ho) and dropped == dro
