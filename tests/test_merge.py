"""convect_shallow_tend's merge after the scheme (uwshcu_merge_shallow_device).

The CPU tests pin the numpy restatement (oracle/merge.py) with hand-computed values
and check the C ABI without a GPU; the GPU test compares the device merge with the
restatement bit for bit on random fields."""
import ctypes as C

import numpy as np
import pytest

import merge as mref
from uwshcu_amd import _capi


def _small():
    pcols, pver = 3, 3
    f = mref.random_fields(pcols, pver, seed=5)
    # cmfmc2 (umf): column 0 convects at the top interface only, column 1 nowhere
    f["umf"][:] = 0.0
    f["umf"][0, 0] = 0.02
    f["umf"][pver, 1] = 0.5          # the surface interface is outside 1..pver: freqsh stays 1
    f["cnt"][:] = [5.0, 2.0, 1.0]
    f["cnt2"][:] = [3.0, 3.0, 1.0]
    f["cnb"][:] = [1.0, 3.0, 2.0]
    f["cnb2"][:] = [2.0, 1.0, 3.0]
    return f, pcols, pver


def test_merge_restatement_known_values():
    f, pcols, pver = _small()
    o = mref.merge(f, ncol=2, pver=pver, latvap=2.5e6)
    np.testing.assert_array_equal(o["freqsh"][:2], [0.0, 1.0])
    np.testing.assert_array_equal(o["cnt"][:2], [3.0, 2.0])          # min(cnt, cnt2)
    np.testing.assert_array_equal(o["cnb"][:2], [2.0, 3.0])          # max(cnb, cnb2)
    assert o["pcnt"][0] == f["pmid"][2, 0] and o["pcnt"][1] == f["pmid"][1, 1]    # pmid(i, int(cnt))
    assert o["pcnb"][0] == f["pmid"][1, 0] and o["pcnb"][1] == f["pmid"][2, 1]
    assert o["cmflq"][1, 0] == f["qtflx"][1, 0] * 2.5e6
    assert o["rprdsh"][2, 1] == f["qrten"][2, 1] + f["qsten"][2, 1]
    assert o["rprdtot"][2, 1] == o["rprdsh"][2, 1] + f["rprddp"][2, 1]
    assert o["cmfmc"][0, 0] == f["cmfmc"][0, 0] + 0.02 and o["cmfmc"][pver, 1] == f["cmfmc"][pver, 1]
    assert o["qc"][0, 1] == f["qc"][0, 1] + f["qc2"][0, 1] and o["rliq"][1] == f["rliq"][1] + f["rliq2"][1]
    np.testing.assert_array_equal(o["massflx_sh"], f["cbmf"])        # all pcols, ncol notwithstanding
    for k in ("cnt", "cnb", "qc", "icwmr", "pcnt"):                 # column 2 >= ncol: untouched
        np.testing.assert_array_equal(o[k][..., 2], f[k][..., 2], err_msg=k)
    assert f["freqsh"][2] != 0.0 and o["freqsh"][2] == 0.0           # :790 freqsh(:) = 0 over all pcols


def test_merge_abi_without_a_gpu():
    assert C.sizeof(_capi.MergeFields) == 8 * len(_capi.MERGE_FIELDS) == 8 * 27
    lib = _capi.lib()
    assert lib.uwshcu_merge_shallow_device(10, 10, 30, C.c_double(2.5e6), None, None) == -10
    assert lib.uwshcu_merge_shallow_device(10, 11, 30, C.c_double(2.5e6), C.byref(_capi.MergeFields()), None) == -10
    assert lib.uwshcu_merge_shallow_device(10, 10, 30, C.c_double(2.5e6), C.byref(_capi.MergeFields()), None) == -10


@pytest.mark.gpu
def test_device_merge_equals_the_restatement():
    import torch
    from uwshcu_amd import device
    pcols, ncol, pver = 5000, 4900, 35
    f = mref.random_fields(pcols, pver, seed=8)
    latvap = _capi.default_consts().xlv
    want = mref.merge(f, ncol, pver, latvap)
    d = {k: torch.from_numpy(v).to("cuda") for k, v in f.items()}
    device.merge_shallow(d, ncol=ncol, latvap=latvap)
    torch.cuda.synchronize()
    for k, v in want.items():
        np.testing.assert_array_equal(d[k].cpu().numpy(), v, err_msg=k)
    # padding columns: freqsh zeroed (convect_shallow.F90:790), massflx_sh = cbmf
    assert np.all(f["freqsh"][ncol:] != 0.0)
    np.testing.assert_array_equal(d["freqsh"][ncol:].cpu().numpy(), 0.0)
