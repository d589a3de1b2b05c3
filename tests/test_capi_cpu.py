"""CPU checks of the C ABI library: it loads and exports every function that
include/sedcol.h declares, and the host structs match the header layout.
No compute calls (there is no GPU here)."""
import ctypes
import re
from pathlib import Path

from sedcol import capi

HEADER = Path(__file__).resolve().parents[1] / "include" / "sedcol.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sedcol_[a-z_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_struct_sizes():
    assert ctypes.sizeof(capi.Params) == 21 * 8 + 7 * 4 + 4  # 7 int32 fields, 4 bytes of tail padding
    assert ctypes.sizeof(capi.ColDiag) == 9 * 8 + 2 * 4
    assert ctypes.sizeof(capi.Desc) == 4 * 4 + 2 * 8
    assert ctypes.sizeof(capi.Schedule) == 4 * 4
    assert ctypes.sizeof(capi.Tuning) == 9 * 4


def test_create_without_device_fails_cleanly():
    import numpy as np
    import torch

    if torch.cuda.is_available():
        return
    try:
        capi.SedimentColumns(np.full(2, 10))
    except capi.SedcolError as e:
        assert e.code == capi.SEDCOL_EHIP
    else:
        raise AssertionError("handle creation succeeded without a device")


def test_descriptor_errors_need_no_device():
    """Descriptor checks run before any HIP call: an empty domain, layer
    counts outside 1..64 and a ragged Fortran layout are argument errors."""
    import numpy as np

    for nl in (np.zeros(0, np.int32), np.array([0, 5]), np.array([65]), np.array([3, 70])):
        try:
            capi.SedimentColumns(nl)
        except capi.SedcolError as e:
            assert e.code == capi.SEDCOL_EINVAL, (nl, e)
        else:
            raise AssertionError(f"nlayers {nl.tolist()} accepted")
    for nl, layout in ((np.array([3, 4]), "fortran"), (np.array([3, 3]), "columns")):
        try:
            capi.SedimentColumns(nl, layout=layout)
        except capi.SedcolError as e:
            assert e.code == capi.SEDCOL_EINVAL, (layout, e)
        else:
            raise AssertionError(f"layout {layout} with nlayers {nl.tolist()} accepted")


def test_warm_start_entry_points_check_arguments_without_a_device():
    """sedcol_set_state_profiles and sedsurr_predict_device reject a null
    handle before any HIP call."""
    import ctypes

    from sedcol import surrogate

    rc = capi.lib().sedcol_set_state_profiles(None, None, 330, 30, None, None, None, 1e-3, None)
    assert rc == capi.SEDCOL_EINVAL
    out = ctypes.c_void_p()
    rc = surrogate.lib().sedsurr_predict_device(None, None, 1, ctypes.byref(out))
    assert rc == capi.SEDCOL_EINVAL and not out.value


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asking for more exits with status 2 and
    says so, before touching any device."""
    import subprocess
    import sys
    from pathlib import Path

    repo = Path(__file__).resolve().parent.parent
    r = subprocess.run([sys.executable, str(repo / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
