"""Every host-side error code of the C ABI, forced through the Python facade
(which calls the C ABI) and checked for its code and message.  The codes
mirror the reference's UG_THROW cases:

- FUNCTION_COUNT  navier_stokes_fv1.cpp:66-68 (constructor, wrong number of functions)
- TRIAL_SPACE     navier_stokes_fv1.cpp:96-103 (prepare_setting: not Lagrange P1)
- HANGING_NODES   navier_stokes_fv1.cpp:92-93, 1528 (non-regular grid)
- UNKNOWN_NAME    upwind_interface.cpp:60-61 (upwind), stabilization.cpp:55-56
                  (stabilisation), stabilization.cpp:97-99 (diffusion length)
- TIME_SERIES     navier_stokes_fv1.cpp:272-274, 621-623 (time-dependent
                  stabilisation without a valid previous time point)
- NO_MESH, ARGUMENT (no reference counterpart: assembly before a mesh, bad
  batched-instance ranges)."""
from __future__ import annotations

import numpy as np
import pytest

from plugin_navierstokes_amd import F64, NSFV1Error, NavierStokesFV1
from plugin_navierstokes_amd import meshes as M

pytestmark = pytest.mark.gpu


def _raises(name, fn, *args, match=None):
    with pytest.raises(NSFV1Error) as e:
        fn(*args)
    assert e.value.name == name, str(e.value)
    if match:
        assert match in str(e.value), str(e.value)
    return e.value


def test_function_count(cuda):
    _raises("FUNCTION_COUNT", NavierStokesFV1, "u,v,w,p", "Inner", 2, match="needs exactly 3 symbolic function")
    _raises("FUNCTION_COUNT", NavierStokesFV1, "u,p", "Inner", 3, match="needs exactly 4 symbolic function")


def test_trial_space_and_hanging_nodes(cuda):
    disc = NavierStokesFV1("u,v,p", "Inner")
    _raises("TRIAL_SPACE", disc.prepare_setting, ["lagrange"] * 2, [1, 1], match="Need exactly 3 functions")
    _raises("TRIAL_SPACE", disc.prepare_setting, ["lagrange", "lagrange", "lagrange"], [1, 2, 1],
            match="Lagrange P1")
    _raises("TRIAL_SPACE", disc.prepare_setting, ["crouzeix-raviart"] * 3, [1, 1, 1], match="Lagrange P1")
    _raises("HANGING_NODES", disc.prepare_setting, ["lagrange"] * 3, [1, 1, 1], True, match="only regular grid")
    disc.prepare_setting(["lagrange"] * 3, [1, 1, 1])


def test_unknown_names(cuda):
    disc = NavierStokesFV1("u,v,p", "Inner")
    _raises("UNKNOWN_NAME", disc.set_upwind, "weighted", match="upwind type 'weighted' not found")
    _raises("UNKNOWN_NAME", disc.set_upwind, "upstream", match="Options are: no, full")
    _raises("UNKNOWN_NAME", disc.set_stabilization, "supg", match="not a valid name of a Schneider-Raw stabilization")
    _raises("UNKNOWN_NAME", disc.set_stabilization, "fields", "cubic",
            match="Diffusion Length calculation method not found")
    disc.set_stabilization("fields")
    _raises("UNKNOWN_NAME", disc.set_stabilization_upwind, "weighted", match="not found")
    # the accepted spellings: trimmed and case-insensitive
    disc.set_upwind("  LinearProfileSkewed ")
    disc.set_stabilization(" FIELDS", "Fivepoint ")


def test_time_series_and_no_mesh(cuda):
    import torch

    mesh = M.cavity_2d(4, M.QUAD, 0.1, 1)
    disc = NavierStokesFV1("u,v,p", "Inner")
    disc.set_stabilization("fields")
    disc.set_upwind("full")
    disc.set_kinematic_viscosity(0.01)
    u = torch.zeros(mesh.nv, 3, dtype=torch.float64, device=cuda)
    _raises("NO_MESH", disc.assemble, u)
    disc.set_mesh(mesh)
    _raises("TIME_SERIES", disc.assemble, u, u, 0.0, match="time step must be positive")
    _raises("TIME_SERIES", disc.assemble, u, u, -0.01, match="time step must be positive")
    disc.assemble(u, u, 0.01, jprec=F64)


def test_batch_offsets_rejected_before_any_write(cuda):
    """Offsets that jump past ne and come back must be refused as ARGUMENT
    (checked before anything is written), and clearing the instances puts
    the handle back on its global viscosity, or on none."""
    import torch

    mesh = M.cavity_2d(4, M.QUAD, 0.1, 1)
    disc = NavierStokesFV1("u,v,p", "Inner")
    disc.set_stabilization("fields")
    disc.set_upwind("full")
    disc.set_mesh(mesh)
    ne = mesh.ne
    _raises("ARGUMENT", disc.batch_set_instances, np.array([0, ne + 1000, ne]), np.array([0.1, 0.2]))
    _raises("ARGUMENT", disc.batch_set_instances, np.array([0, 3, 2, ne]), np.array([0.1, 0.2, 0.3]))
    _raises("ARGUMENT", disc.batch_set_instances, np.array([0, ne]), np.array([-1.0]))
    u = torch.zeros(mesh.nv, 3, dtype=torch.float64, device=cuda)
    disc.batch_set_instances(np.array([0, ne // 2, ne]), np.array([0.1, 0.2]))
    disc.assemble(u)
    disc.batch_set_instances(np.zeros(1, dtype=np.int64), np.zeros(0))
    _raises("NO_VISCOSITY", disc.assemble, u)
    disc.set_kinematic_viscosity(0.05)
    disc.assemble(u)
