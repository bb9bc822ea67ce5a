"""Contraction DAG plans (jet_hip_dag_estimate, no device): the forward
contraction of a triangle network A(a,b) B(b,c) C(c,a) and the environments
of its leaves, with shared intermediates.

* flops: the sum of 8*M*N*K over the steps, worked out by hand below;
* arena: the working region (two live 256-byte intermediates at most) plus
  the kept region (four results of 256 bytes each after rounding);
* errors: an operand that does not exist yet, a kept leaf, a result kept
  twice.
"""
import numpy as np
import pytest

from jet_amd import _native as nat
from jet_amd.network import dag_estimate

INDICES = [["a", "b"], ["b", "c"], ["c", "a"], []]
SHAPES = [[2, 3], [3, 4], [4, 2], []]
# nodes: 0 A, 1 B, 2 C, 3 one; 4 = AB, 5 = ABC, 6 = E_AB, 7 = E_C, 8 = E_A, 9 = E_B
STEPS = [(0, 1), (4, 2), (3, 2), (3, 4), (6, 1), (6, 0)]
KEEP = [5, 8, 9, 7]


def test_dag_flops_and_arena():
    flops, arena, ws = dag_estimate(INDICES, SHAPES, STEPS, KEEP, np.complex64)
    # AB: M=a=2, N=c=4, K=b=3; ABC: K=a*c=8; E_AB = one.C: N=8; E_C = one.AB: N=8;
    # E_A = E_AB(c,a).B(b,c): M=a=2, N=b=3, K=c=4; E_B = E_AB.A(a,b): M=c=4, N=b=3, K=a=2.
    assert flops == 8 * (2 * 4 * 3 + 8 + 8 + 8 + 2 * 3 * 4 + 4 * 3 * 2)
    # working region: AB and E_AB live together (2 x 256 B); kept: 4 x 256 B
    assert arena == 2 * 256 + 4 * 256
    assert ws >= 0


def test_dag_keeps_everything_it_is_asked_to():
    # keeping AB too moves it out of the working region
    _, arena, _ = dag_estimate(INDICES, SHAPES, STEPS, KEEP + [4], np.complex64)
    assert arena == 256 + 5 * 256


@pytest.mark.parametrize("steps,keep,msg", [
    ([(0, 1), (5, 2)], [4], "Node ID in contraction path pair is invalid."),
    ([(0, 1), (4, 2)], [2], "Only step results can be kept."),
    ([(0, 1), (4, 2)], [5, 5], "A step result is kept twice."),
    ([(0, 0)], [], "Node ID in contraction path pair is invalid."),
])
def test_dag_errors(steps, keep, msg):
    with pytest.raises(nat.JetHipError, match=msg):
        dag_estimate(INDICES, SHAPES, steps, keep, np.complex64)
