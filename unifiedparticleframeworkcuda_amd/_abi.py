"""ctypes mirror of the C ABI structs in include/upf.h.

The field order must match the C declarations exactly; tests/test_abi.py
checks sizes and offsets against the compiled library's own report.
"""
import ctypes as C

UPF_OK = 0
UPF_ERR_HIP = 1
UPF_ERR_ARG = 2
UPF_ERR_RANGE = 3
UPF_ERR_NAN = 4
UPF_ERR_NBR_OVERFLOW = 5
UPF_ERR_DELTA = 6
UPF_ERR_NO_DEVICE = 7
UPF_ERR_ISOLATED = 8
UPF_ERR_STATE = 9
UPF_ERR_IO = 10

ERROR_NAMES = {
    0: "UPF_OK", 1: "UPF_ERR_HIP", 2: "UPF_ERR_ARG", 3: "UPF_ERR_RANGE",
    4: "UPF_ERR_NAN", 5: "UPF_ERR_NBR_OVERFLOW", 6: "UPF_ERR_DELTA",
    7: "UPF_ERR_NO_DEVICE", 8: "UPF_ERR_ISOLATED", 9: "UPF_ERR_STATE", 10: "UPF_ERR_IO",
}

UPF_RIGID = 0
UPF_LIQUID = 2
UPF_FROZEN = 6

UPF_TEXT_LIQUID = 1
UPF_TEXT_RIGID = 2
UPF_TEXT_POSITIONS = 3
UPF_TEXT_RESTART = 4

UPF_PCISPH_STATIC = 0
UPF_WCSPH_STATIC = 1
UPF_PCISPH_RIGID = 2
UPF_SPH_STATIC = 3

UPF_BOUNDARY_CPU_REF = 0
UPF_BOUNDARY_AKINCI = 1

F3 = C.c_float * 3


class Params(C.Structure):
    _fields_ = [
        ("spacing", C.c_float),
        ("particle_radius", C.c_float),
        ("support_radius", C.c_float),
        ("force_distance", C.c_float),
        ("initial_mass", C.c_float),
        ("rest_density", C.c_float),
        ("visc_const", C.c_float),
        ("gravity", C.c_float),
        ("dt", C.c_float),
        ("gas_const_wcsph", C.c_float),
        ("gamma", C.c_float),
        ("max_boundary_force", C.c_float),
        ("add_boundary_force", C.c_int),
        ("density_error_factor", C.c_float),
        ("density_error_factor_parameter", C.c_float),
        ("max_density_error_allowed", C.c_float),
        ("min_loops", C.c_int),
        ("max_loops", C.c_int),
        ("kernel_self", C.c_float),
        ("surface_tension_gamma", C.c_float),
        ("surface_adhesion_beta", C.c_float),
        ("spring_coefficient", C.c_float),
        ("spring_coefficient_boundary", C.c_float),
        ("damping_coefficient", C.c_float),
        ("terminal_speed", C.c_float),
        ("boundary_pressure_mode", C.c_int),
        ("renormalize_quaternion", C.c_int),
        ("rigid_predicted_current", C.c_int),
        ("vbox_min", F3),
        ("vbox_max", F3),
        ("scales", F3),
        ("block_size", C.c_uint32),
        ("block_shift", C.c_uint32),
        ("grid_dim", C.c_uint32),
        ("collision_min", F3),
        ("collision_max", F3),
        ("support_radius_sq_limit", C.c_float),
        ("max_neighbors", C.c_int),
        ("add_wall_weight", C.c_int),
        ("container_min", F3),
        ("container_max", F3),
        ("gas_const_sph", C.c_float),
    ]

    def copy(self):
        p = Params()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(Params))
        return p

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class DamBreakDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("box", F3),
        ("jitter_rel", C.c_float),
        ("wall_rel", C.c_float),
        ("lattice_rel", C.c_float),
        ("seed", C.c_uint64),
    ]


class RigidSceneDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("box", F3),
        ("n_bodies", C.c_int),
        ("size_min", C.c_float), ("size_max", C.c_float),
        ("density_min", C.c_float), ("density_max", C.c_float),
        ("jitter_rel", C.c_float),
        ("surface_rel", C.c_float),
        ("wall_rel", C.c_float),
        ("lattice_rel", C.c_float),
        ("seed", C.c_uint64),
    ]


class Body(C.Structure):
    _fields_ = [
        ("pos", F3), ("mass", C.c_float),
        ("vel", F3), ("pad0", C.c_float),
        ("angular_velocity", F3), ("pad1", C.c_float),
        ("linear_momentum", F3), ("pad2", C.c_float),
        ("angular_momentum", F3), ("pad3", C.c_float),
        ("quaternion", C.c_float * 4),
        ("rotation", C.c_float * 9),
        ("inv_inertia_local", C.c_float * 9),
        ("inv_inertia_world", C.c_float * 9),
        ("force", F3),
        ("torque", F3),
    ]


class Tuning(C.Structure):
    """upf_tuning: launch shapes that change speed only."""
    _fields_ = [("loop_block", C.c_int), ("grid_cap", C.c_int), ("nbr_rounds", C.c_int),
                ("nbr_stage", C.c_int), ("speculate", C.c_int), ("loop_list", C.c_int), ("loop_width", C.c_int),
                ("nbr_tile", C.c_int), ("spec_first", C.c_int), ("xcd_remap", C.c_int), ("nbr_skip", C.c_int),
                ("nbr_fuse", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}

    @classmethod
    def from_dict(cls, d):
        t = cls()
        for k, _ in cls._fields_:
            setattr(t, k, int(d[k]))
        return t


def ptr(a, ctype=C.c_float):
    """Pointer to a numpy array's data (None passes NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))
