"""Host-side checks of the Python binding (no GPU): output arrays the library
writes into are validated before any call crosses the C ABI, and a mobility
handle is released as soon as its last reference goes (no reference cycle
between the mobility and its operator view).  The library is replaced by a
recording stand-in, so no device call is made."""
import gc

import numpy as np
import pytest

import spheresimulator_amd as S
from spheresimulator_amd import bie


class _FakeLib:
    """Records the calls the binding makes; every call succeeds."""

    def __init__(self, n_dof=12):
        self.calls = []
        self.n_dof = n_dof

    def __getattr__(self, name):
        def f(*args):
            self.calls.append(name)
            if name == "sphstk_operator_sizes":
                args[1]._obj.value = self.n_dof // 9
                args[2]._obj.value = self.n_dof // 3
                args[3]._obj.value = self.n_dof
                args[4]._obj.value = 0
            if name == "sphstk_mobility_operator":
                return 1234
            if name.endswith("_create_batched"):
                args[-1]._obj.value = 77
            return 0
        return f


@pytest.fixture
def fake(monkeypatch):
    lib = _FakeLib()
    monkeypatch.setattr(bie, "load", lambda: lib)
    return lib


def _operator(fake):
    return S.SphereSTKSHOperator(None, None, None, None, _handle=99)


@pytest.mark.parametrize("bad", ["float32", "short", "reversed", "readonly", "list"])
def test_output_arrays_are_validated(fake, bad):
    op = _operator(fake)
    n = op.n_dof
    x = np.ones(n)
    y = {"float32": np.zeros(n, dtype=np.float32), "short": np.zeros(n - 1), "reversed": np.zeros(n)[::-1],
         "readonly": np.zeros(n), "list": [0.0] * n}[bad]
    if bad == "readonly":
        y.flags.writeable = False
    for call in (lambda: op.apply(x, y), lambda: op.apply_inner(x, y), lambda: op.apply_outer(x, y),
                 lambda: op.applyP2POP(x, y, 1.0, 0.0, 0.0), lambda: op.applyLOP(x, y, 1.0),
                 lambda: op.projectNullSpace(y)):
        with pytest.raises(ValueError):
            call()
    assert not any(c.startswith("sphstk_operator_apply") or c == "sphstk_operator_project_nullspace"
                   for c in fake.calls)


def test_valid_output_arrays_reach_the_library(fake):
    op = _operator(fake)
    n = op.n_dof
    op.apply(np.ones(n), np.zeros(n))
    op.applyP2POP(np.ones(n), np.zeros(n), 1.0, 0.0, 0.0)
    op.projectNullSpace(np.zeros(n))
    assert fake.calls.count("sphstk_operator_apply") == 1
    assert fake.calls.count("sphstk_operator_apply_p2p") == 1
    assert fake.calls.count("sphstk_operator_project_nullspace") == 1


def test_mobility_apply_validates_Y(fake):
    M = S.SphereSTKMobMat(np.zeros((1, 3)), [1.0], None, 4)
    X = np.ones(6)
    for Y in (np.zeros(6, dtype=np.float32), np.zeros(5), np.zeros(12)[::2]):
        with pytest.raises(ValueError):
            M.apply(X, Y)
    assert "sphstk_mobility_apply" not in fake.calls
    M.apply(X, np.zeros(6))
    assert "sphstk_mobility_apply" in fake.calls


def test_mobility_is_released_without_garbage_collection(fake):
    gc.disable()
    try:
        M = S.SphereSTKMobMat(np.zeros((1, 3)), [1.0], None, 4)
        view = M.op
        assert view.n_dof == fake.n_dof
        del view
        assert "sphstk_mobility_destroy" not in fake.calls
        del M
        assert fake.calls.count("sphstk_mobility_destroy") == 1
        # an operator view keeps its mobility alive until the view goes
        M = S.SphereSTKMobMat(np.zeros((1, 3)), [1.0], None, 4)
        view = M.op
        del M
        assert fake.calls.count("sphstk_mobility_destroy") == 1
        del view
        assert fake.calls.count("sphstk_mobility_destroy") == 2
        # the view never destroys the operator it does not own
        assert "sphstk_operator_destroy" not in fake.calls
    finally:
        gc.enable()
