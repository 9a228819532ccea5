"""Helpers that give a problem non-zero Dirichlet (fixed-DOF) values.

The reference keeps the FIRST constraint that holds an entity: the
constraint loop stops at the first match (Get_ConstraintOfElement.cpp:51-54)
and Dof_DefineAssignFixedDof never re-assigns a fixed DOF
(DofData.cpp:1137-1151).  ``with_dirichlet`` therefore appends duplicates of
some constrained entities with different values; a last-wins rule would take
those instead and change the right-hand side.
"""
import numpy as np

from getdp_amd import problems as P


def with_dirichlet(pr: P.Problem, seed: int = 0, ndup: int = 16) -> P.Problem:
    rng = np.random.default_rng(seed)
    for s in pr.spaces:
        ent = np.asarray(s.fixed_entity, np.int32)
        if ent.size == 0:
            continue
        val = rng.uniform(-2.0, 3.0, size=ent.size)
        k = min(ndup, ent.size)
        dup = rng.choice(ent.size, size=k, replace=False)
        # duplicates both after (a second group holding a corner) and before
        # the first group: the earliest occurrence must win in both cases
        head = dup[: k // 2]
        tail = dup[k // 2:]
        s.fixed_entity = np.concatenate([ent[head], ent, ent[tail]]).astype(np.int32)
        s.fixed_value = np.concatenate([rng.uniform(10, 20, size=head.size), val,
                                        rng.uniform(-20, -10, size=tail.size)])
    return pr


def first_values(space: P.Space) -> dict:
    """entity -> value of its first occurrence in the constraint list."""
    out = {}
    for e, v in zip(space.fixed_entity.tolist(), space.fixed_value.tolist()):
        out.setdefault(e, v)
    return out
