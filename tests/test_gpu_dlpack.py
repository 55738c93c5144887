"""Zero-copy state exchange with torch on the device: DLPack export of the
state (lq_sv_dlpack), device-to-device updateData / getState
(lq_sv_update_data_device / lq_sv_get_data_device), and the lifetime of an
exported state. Values are compared bit for bit with the host copies of the
same state (and with the oracle for the circuit run on it, 1e-12)."""
import gc

import numpy as np
import pytest

from helpers import oracle_run, rand_state

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
L = pytest.importorskip("lightning_amd.lightning_qubit_ops")
W = pytest.importorskip("lightning_amd.workloads")


def make(bits, n):
    return (L.StateVectorC128 if bits == 128 else L.StateVectorC64)(n)


@pytest.mark.parametrize("bits", [128, 64])
def test_export_shares_the_state(bits):
    n = 12
    sv = make(bits, n)
    ops = W.hea_rot_cz(n, 2, seed=7)
    W.apply_circuit(sv, ops)
    t = sv.to_torch()
    assert t.device.type == "cuda" and t.shape == (1 << n,)
    assert t.dtype == (torch.complex128 if bits == 128 else torch.complex64)
    assert t.data_ptr() == sv.device_ptr()
    host = sv.state()
    np.testing.assert_array_equal(t.cpu().numpy(), host)
    np.testing.assert_allclose(host, oracle_run(n, ops), atol=1e-12 if bits == 128 else 5e-6)
    # gates applied after the export show in the tensor
    sv.PauliX([3], False, [])
    sv.synchronize()
    np.testing.assert_array_equal(t.cpu().numpy(), sv.state())
    # torch.from_dlpack on the state object itself
    t2 = torch.from_dlpack(sv)
    assert t2.data_ptr() == t.data_ptr()


def test_export_outlives_the_state_object():
    n = 10
    sv = make(128, n)
    W.apply_circuit(sv, W.hea_rot_cz(n, 1, seed=3))
    want = sv.state()
    t = sv.to_torch()
    del sv
    gc.collect()
    other = make(128, n)      # allocations after the release must not reuse the exported memory
    other.PauliX([0], False, [])
    other.synchronize()
    np.testing.assert_array_equal(t.cpu().numpy(), want)
    del t
    gc.collect()


def test_export_of_a_deferred_reset_and_an_unconsumed_capsule():
    sv = make(128, 9)
    t = sv.to_torch()          # the reset is written before the export
    want = np.zeros(1 << 9, np.complex128)
    want[0] = 1.0
    np.testing.assert_array_equal(t.cpu().numpy(), want)
    cap = sv.__dlpack__()      # never consumed: released with the capsule
    del cap
    gc.collect()
    assert sv.__dlpack_device__()[0] == 10


@pytest.mark.parametrize("bits", [128, 64])
def test_device_update_and_get(bits):
    n = 11
    rng = np.random.default_rng(4)
    psi = rand_state(n, rng).astype(np.complex128 if bits == 128 else np.complex64)
    src = torch.from_numpy(psi).to("cuda")
    sv = make(bits, n)
    sv.updateData(src)
    np.testing.assert_array_equal(sv.state(), psi)
    W.apply_circuit(sv, W.hea_rot_cz(n, 1, seed=5))
    dst = torch.zeros(1 << n, dtype=src.dtype, device="cuda")
    sv.getState(dst)
    np.testing.assert_array_equal(dst.cpu().numpy(), sv.state())
    np.testing.assert_allclose(sv.state(), oracle_run(n, W.hea_rot_cz(n, 1, seed=5), psi),
                               atol=1e-12 if bits == 128 else 5e-6)
    # an unaligned source (a view one amplitude in) goes through the runtime's copy
    big = torch.from_numpy(np.concatenate([[0], psi]).astype(psi.dtype)).to("cuda")
    sv.updateData(big[1:])
    np.testing.assert_array_equal(sv.state(), psi)
    with pytest.raises(L.LightningException):
        sv.updateData(src[:-1].contiguous())
    with pytest.raises(ValueError):
        sv.getState(torch.zeros(1 << n, dtype=torch.complex128 if bits == 64 else torch.complex64, device="cuda"))
