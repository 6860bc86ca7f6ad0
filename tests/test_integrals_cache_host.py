"""The integral API's reuse of small GPU bases (quiqbox_amd.integrals._basis):
the same functions give the same basis object, any change of a centre,
exponent, coefficient or angular part gives a new one, and large bases are
never kept.  GPUBasis is replaced by a counting stand-in, so no GPU is needed."""
import pytest

import quiqbox_amd as qb
from quiqbox_amd import integrals


class _Stub:
    made = 0

    def __init__(self, flat):
        _Stub.made += 1
        self.n = flat.n_bf


@pytest.fixture
def stub(monkeypatch):
    monkeypatch.setattr(integrals, "GPUBasis", _Stub)
    monkeypatch.setattr(integrals, "_cache", integrals.collections.OrderedDict())
    _Stub.made = 0
    return _Stub


def _orb(x=0.0, a=1.2, c=0.7, ijk=(0, 0, 0)):
    return qb.genGaussTypeOrb((x, 0.1, -0.2), [a, 0.4], [c, 0.5], ijk)


def test_same_functions_reuse_one_basis(stub):
    o1, o2 = _orb(), _orb(1.0)
    b1 = integrals._basis([o1, o2])
    b2 = integrals._basis([o1, o2])
    b3 = integrals._basis([_orb(), _orb(1.0)])      # equal content, new objects
    assert b1 is b2 is b3 and stub.made == 1


@pytest.mark.parametrize("change", [dict(x=0.5), dict(a=1.3), dict(c=0.6), dict(ijk=(1, 0, 0))])
def test_any_parameter_change_gives_a_new_basis(stub, change):
    base = integrals._basis([_orb(), _orb(1.0)])
    other = integrals._basis([_orb(**change), _orb(1.0)])
    assert other is not base and stub.made == 2


def test_order_matters(stub):
    o1, o2 = _orb(), _orb(1.0)
    assert integrals._basis([o1, o2]) is not integrals._basis([o2, o1])


def test_large_bases_are_not_kept(stub):
    orbs = [_orb(0.1 * k) for k in range(integrals._CACHE_MAX_BF + 1)]
    integrals._basis(orbs)
    integrals._basis(orbs)
    assert stub.made == 2 and len(integrals._cache) == 0


def test_oldest_entry_dropped(stub):
    for k in range(integrals._CACHE_MAX + 3):
        integrals._basis([_orb(0.01 * k)])
    assert len(integrals._cache) == integrals._CACHE_MAX
    integrals._basis([_orb(0.0)])                 # the first one was dropped
    assert stub.made == integrals._CACHE_MAX + 4
