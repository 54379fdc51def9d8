"""CPU: the circuit layer's host logic through the C ABI -- levelisation, NOT/COPY folding (no
extra level), wire bookkeeping and argument errors.  No device is touched."""
import pytest

import tfhe_amd as T


def test_ripple_carry_adder_levels():
    c = T.ripple_carry_adder(32)
    info = c.info()
    # one wide level of 32 XOR, then one MUX (two bootstraps) + one XOR per bit along the carry chain
    assert info == {"levels": 33, "max_level_bootstraps": 32, "bootstraps": 32 + 32 * 3, "outputs": 33}
    for i in range(32):
        assert c.level_of(64 + i) == 1          # p_i
    c.free()


def test_adder_widths():
    for bits in (1, 4, 8):
        c = T.ripple_carry_adder(bits)
        assert c.info()["levels"] == bits + 1
        assert c.info()["bootstraps"] == 4 * bits
        c.free()


def test_not_and_copy_do_not_add_levels():
    c = T.Circuit(2)
    x = c.add("AND", 0, 1)
    y = c.add("NOT", x)
    z = c.add("COPY", y)
    w = c.add("OR", z, 0)
    k = c.constant(1)
    nk = c.add("NOT", k)
    assert [c.level_of(v) for v in (0, 1, x, y, z, w, k, nk)] == [0, 0, 1, 1, 1, 2, 0, 0]
    c.output(w)
    c.output(nk)
    assert c.info() == {"levels": 2, "max_level_bootstraps": 1, "bootstraps": 2, "outputs": 2}
    c.free()


def test_mux_counts_two_bootstraps():
    c = T.Circuit(3)
    m = c.add("MUX", 0, 1, 2)
    c.output(m)
    assert c.info()["bootstraps"] == 2 and c.info()["levels"] == 1
    c.free()


@pytest.mark.parametrize("bad", [
    lambda c: c.add("AND", 0, 5),
    lambda c: c.add("AND", -1, 0),
    lambda c: c.add("MUX", 0, 1, 9),
    lambda c: c.add(99, 0, 1),
    lambda c: c.add("NOT", 7),
    lambda c: c.constant(2),
    lambda c: c.output(42),
    lambda c: c.level_of(42),
])
def test_bad_arguments_raise(bad):
    c = T.Circuit(2)
    with pytest.raises(T.TfheError):
        bad(c)
    c.free()


def test_negative_inputs_rejected():
    with pytest.raises(T.TfheError):
        T.Circuit(-1)
