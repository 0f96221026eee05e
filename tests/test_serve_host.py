"""Admission projection of the served loop (fsig_project_bytes, host only):
closed forms on one-vertex, one-edge and triangle queries, and the dead
query of an absent label."""
import numpy as np

import fsig_amd as F

N, L = 1000, 3
LC = np.array([100, 300, 600], np.uint64)
PE = np.array([[40, 70, 90], [70, 500, 800], [90, 800, 2000]], np.uint64)


def test_single_vertex():
    q = F.Query(1, [1], np.zeros((0, 2)))
    assert F.project_bytes(N, LC, PE, q) == 4 * N + 9 * 300 + 512


def test_edge_query_smaller_class_first():
    # labels 0 and 2: candidate tables of both classes, the candidate-edge
    # lists of both directions, one stored level (100 rows of 1 column + 12)
    q = F.Query(2, [2, 0], [(0, 1)])
    want = (4 * N + (9 * 100 + 512) + (9 * 600 + 512) + 4 * (100 + 1) + 4 * 90 + 4 * (600 + 1) + 4 * 90
            + 100 * 16)
    assert F.project_bytes(N, LC, PE, q) == want


def test_triangle_one_class():
    # one class; level 1 = 300 rows (16 B each), level 2 = 300 * (500 / 300)
    # = 500 rows (20 B each); the last level is counted, not stored
    q = F.Query(3, [1, 1, 1], [(0, 1), (1, 2), (0, 2)])
    want = 4 * N + (9 * 300 + 512) + 4 * (300 + 1) + 4 * 500 + 300 * 16 + 500 * 20
    assert F.project_bytes(N, LC, PE, q) == want


def test_absent_label_is_dead():
    q = F.Query(2, [0, 7], [(0, 1)])
    assert F.project_bytes(N, LC, PE, q) == 4096


def test_denser_query_projects_fewer_rows():
    # a 4-cycle with a chord keeps fewer level-3 rows than the 4-path with
    # the same labels (the extra backward neighbour multiplies by an edge
    # probability below 1)
    path = F.Query(4, [1, 1, 1, 1], [(0, 1), (1, 2), (2, 3)])
    chord = F.Query(4, [1, 1, 1, 1], [(0, 1), (1, 2), (2, 3), (3, 0), (0, 2)])
    assert F.project_bytes(N, LC, PE, chord) < F.project_bytes(N, LC, PE, path)
