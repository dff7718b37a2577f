"""The 9-input sorting network of the 3x3 medians (median_kernel, median3_kernel, spike_kernel).

A comparator network sorts every input exactly when it sorts every 0/1 input (Knuth, TAOCP vol. 3,
5.3.4, the zero-one principle), so the 2^9 binary inputs decide it; the network is read from both
kernel sources, which must hold the same one. With fminf/fmaxf comparators and no NaN among the
values (invalid neighbours are +inf), a sorting network yields the same ascending values as the
reference's sort, whatever the network.
"""
import itertools
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOURCES = ["open_source_einstar_amd/csrc/sgm_kernels.hip", "open_source_einstar_amd/csrc/refine_kernels.hip"]


def network(path):
    text = open(os.path.join(ROOT, path)).read()
    m = re.search(r"kNet\[25\]\[2\]\s*=\s*\{(.*?)\};", text, re.S)
    assert m, f"no kNet[25][2] in {path}"
    pairs = [(int(a), int(b)) for a, b in re.findall(r"\{\s*(\d+)\s*,\s*(\d+)\s*\}", m.group(1))]
    assert len(pairs) == 25
    return pairs


def apply(net, values):
    v = list(values)
    for i, j in net:
        v[i], v[j] = min(v[i], v[j]), max(v[i], v[j])
    return v


def test_both_kernels_use_the_same_network():
    assert network(SOURCES[0]) == network(SOURCES[1])


def test_network_sorts_every_zero_one_input():
    net = network(SOURCES[0])
    assert all(0 <= i < j < 9 for i, j in net)
    for bits in itertools.product((0, 1), repeat=9):
        assert apply(net, bits) == sorted(bits), bits


def test_network_sorts_values_with_infinities():
    net = network(SOURCES[0])
    inf = float("inf")
    for values in ([3.5, inf, -1.0, 2.0, inf, 0.5, 7.25, -1.0, 4.0], [inf] * 9, [1.0] * 9,
                   [9, 8, 7, 6, 5, 4, 3, 2, 1]):
        assert apply(net, values) == sorted(values)
