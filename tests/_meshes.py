"""Mesh variants for the tests of quadrilateral/hexahedral meshes that are not
structured boxes: elements relabelled by a rotation of the reference
square/cube (same geometry, other local vertex order, so neighbours meet
through other local faces in other orientations), and a smooth periodic
displacement of the vertices (bilinear/trilinear, non-affine elements).
"""
import itertools

import numpy as np


def rotations(dim, quarter_turns=True):
    """Signed permutation matrices of determinant +1 (the rotations of the
    square / cube); without quarter turns only the diagonal ones."""
    out = []
    for perm in itertools.permutations(range(dim)):
        for signs in itertools.product((1, -1), repeat=dim):
            R = np.zeros((dim, dim))
            for i, j in enumerate(perm):
                R[i, j] = signs[i]
            if np.linalg.det(R) > 0 and (quarter_turns or np.allclose(np.abs(R), np.eye(dim))):
                out.append(R)
    return out


def _corner(v, dim):
    return np.array([2 * ((v >> d) & 1) - 1 for d in range(dim)], dtype=float)


def _vertex(c):
    return sum(int((ci + 1) // 2) << d for d, ci in enumerate(c))


def rotate_elements(X, nodes, seed, quarter_turns=True):
    """Relabel every element's vertices by a random rotation R: new local
    vertex v is the old vertex at corner R c(v), so the new element map is
    x_new(xi) = x_old(R xi). Returns (X, nodes, rotations used)."""
    rng = np.random.default_rng(seed)
    ne, nv, dim = X.shape
    Rs = rotations(dim, quarter_turns)
    X2, N2, used = X.copy(), nodes.copy(), []
    for e in range(ne):
        R = Rs[rng.integers(len(Rs))]
        used.append(R)
        for v in range(nv):
            old = _vertex(R @ _corner(v, dim))
            X2[e, v] = X[e, old]
            N2[e, v] = nodes[e, old]
    return X2, N2, used


def perturb(X, eps=0.04):
    """x_r += eps sin(2 pi x_{dim-1-r}) sin(2 pi x_r): smooth, periodic with
    period 1, zero on the planes x_r = 0, 1 (the box and the identification of
    periodic copies are kept), and a different displacement at each vertex of
    an element, so the elements are no longer parallelograms/parallelepipeds."""
    return X + eps * np.sin(2 * np.pi * X[..., ::-1]) * np.sin(2 * np.pi * X)
