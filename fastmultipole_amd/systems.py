"""User systems with the built-in point kernels, mirroring the reference's test systems.

Gravitational follows test/gravitational.jl (Point{Source}, data_per_body 5,
buffer rows [x, y, z, radius, strength]); VortexParticles follows
test/vortex.jl (Point{Vortex}, data_per_body 7, rows [x, y, z, sigma,
Gamma_x, Gamma_y, Gamma_z]); DipolePoints follows the DipolePoints system of
test/bodytomultipole_test.jl:122-168 (Point{Dipole}, rows [x, y, z, 0, q]);
GaussianVortexParticles is VortexParticles with the Gaussian-core nearfield
kernel (sigma = row 4). Each implements the reference's user-system
interface (src/compatibility.jl): source_system_to_buffer!, data_per_body,
get_position, strength_dims, get_n_bodies, has_vector_potential,
buffer_to_target_system! and the body_to_multipole! kernel tag (fmm_kernel).
"""
import numpy as np

from . import _lib as L


class Gravitational:
    """test/gravitational.jl:16-119. bodies: (N, >=5) rows [x, y, z, radius, strength, ...]."""

    fmm_kernel = L.POINT_SOURCE       # body_to_multipole!(::Gravitational) = Point{Source} (:74)

    def __init__(self, bodies):
        b = np.asarray(bodies, dtype=np.float64)
        self.position = np.ascontiguousarray(b[:, 0:3])
        self.radius = np.ascontiguousarray(b[:, 3])
        self.strength = np.ascontiguousarray(b[:, 4])
        self.potential = np.zeros((b.shape[0], 16))   # columns 0 phi, 4:7 gradient, 7:16 hessian (reference rows 1..16)

    # --- user-system interface (src/compatibility.jl:21-286)
    def get_n_bodies(self):
        return self.position.shape[0]

    def data_per_body(self):
        return 5

    def strength_dims(self):
        return 1

    def has_vector_potential(self):
        return False

    def get_position(self, i):
        return self.position[i]

    def positions(self):
        return self.position

    def source_buffer(self):
        """All bodies through source_system_to_buffer! (:50-54), as an (N, 5) array."""
        return np.column_stack([self.position, self.radius, self.strength])

    def set_strengths(self, strengths):
        """buffer_to_system_strength! (test/solve_test.jl:27-32)."""
        self.strength = np.ascontiguousarray(strengths, dtype=np.float64)

    def buffer_to_target_system(self, target_buffer, switch):
        """buffer_to_target_system! (:106-119): assigns phi, gradient and hessian (zero when switched off)."""
        n = self.get_n_bodies()
        self.potential[:, 0] = target_buffer[:n, 3] if switch.scalar_potential else 0.0
        self.potential[:, 4:7] = target_buffer[:n, 4:7] if switch.gradient else 0.0
        self.potential[:, 7:16] = target_buffer[:n, 7:16] if switch.hessian else 0.0

    def reset(self):
        self.potential[:] = 0.0


class VortexParticles:
    """test/vortex.jl:20-147. position (N,3), strength (N,3), sigma (N,)."""

    fmm_kernel = L.POINT_VORTEX       # body_to_multipole!(::VortexParticles) = Point{Vortex} (:78)

    def __init__(self, position, strength, sigma=None):
        self.position = np.ascontiguousarray(position, dtype=np.float64)
        self.strength = np.ascontiguousarray(strength, dtype=np.float64)
        n = self.position.shape[0]
        self.sigma = np.zeros(n) if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64)
        self.potential = np.zeros((n, 13))            # reference: potential rows 5:13 hold the velocity gradient
        self.gradient_stretching = np.zeros((n, 6))   # rows 1:3 velocity

    def get_n_bodies(self):
        return self.position.shape[0]

    def data_per_body(self):
        return 7

    def strength_dims(self):
        return 3

    def has_vector_potential(self):
        return True

    def get_position(self, i):
        return self.position[i]

    def positions(self):
        return self.position

    def source_buffer(self):
        """source_system_to_buffer! (:56-60): rows [x, y, z, sigma, Gamma]."""
        return np.column_stack([self.position, self.sigma, self.strength])

    def buffer_to_target_system(self, target_buffer, switch):
        """buffer_to_target_system! (:130-143): accumulates velocity and velocity gradient."""
        n = self.get_n_bodies()
        if switch.gradient:
            self.gradient_stretching[:, 0:3] += target_buffer[:n, 4:7]
        if switch.hessian:
            self.potential[:, 4:13] += target_buffer[:n, 7:16]

    def reset(self):
        self.potential[:] = 0.0
        self.gradient_stretching[:] = 0.0


class DipolePoints:
    """test/bodytomultipole_test.jl:122-168: point dipoles of vector strength q (Point{Dipole})."""

    fmm_kernel = L.POINT_DIPOLE

    def __init__(self, position, strength):
        self.position = np.ascontiguousarray(position, dtype=np.float64)
        self.strength = np.ascontiguousarray(strength, dtype=np.float64)
        n = self.position.shape[0]
        self.potential = np.zeros(n)
        self.gradient = np.zeros((n, 3))
        self.hessian = np.zeros((n, 9))

    def get_n_bodies(self):
        return self.position.shape[0]

    def data_per_body(self):
        return 7

    def strength_dims(self):
        return 3

    def has_vector_potential(self):
        return False

    def get_position(self, i):
        return self.position[i]

    def positions(self):
        return self.position

    def source_buffer(self):
        return np.column_stack([self.position, np.zeros(self.get_n_bodies()), self.strength])

    def buffer_to_target_system(self, target_buffer, switch):
        n = self.get_n_bodies()
        if switch.scalar_potential:
            self.potential += target_buffer[:n, 3]
        if switch.gradient:
            self.gradient += target_buffer[:n, 4:7]
        if switch.hessian:
            self.hessian += target_buffer[:n, 7:16]

    def reset(self):
        self.potential[:] = 0.0
        self.gradient[:] = 0.0
        self.hessian[:] = 0.0


class GaussianVortexParticles(VortexParticles):
    """Vortex particles with Gaussian cores of width sigma: the nearfield uses the regularised
    kernel, the expansions are those of Point{Vortex}."""

    fmm_kernel = L.VORTEX_GAUSSIAN


class SourceVortexPoints(VortexParticles):
    """Bodies carrying a point source and a point vortex (Point{SourceVortex}, src/bodytomultipole.jl:491-496):
    position (N,3), source strength q (N,), vortex strength (N,3), sigma (N,). Expansions: the source's then the
    vortex's; nearfield: the point source's terms plus the singular vortex's."""

    fmm_kernel = L.POINT_SOURCE_VORTEX

    def __init__(self, position, source_strength, strength, sigma=None):
        super().__init__(position, strength, sigma)
        self.source_strength = np.ascontiguousarray(source_strength, dtype=np.float64).reshape(-1)
        self.scalar_potential = np.zeros(self.get_n_bodies())

    def data_per_body(self):
        return 8

    def strength_dims(self):
        return 4

    def source_buffer(self):
        """rows [x, y, z, sigma, q, Gamma]"""
        return np.column_stack([self.position, self.sigma, self.source_strength, self.strength])

    def buffer_to_target_system(self, target_buffer, switch):
        n = self.get_n_bodies()
        if switch.scalar_potential:
            self.scalar_potential += target_buffer[:n, 3]
        super().buffer_to_target_system(target_buffer, switch)

    def reset(self):
        super().reset()
        self.scalar_potential[:] = 0.0


class _Panels:
    """Constant-strength planar panels (test/bodytomultipole_test.jl:80-120 SourcePanels,
    :221-272 DipolePanels): vertices (N, 3|4, 3) counter-clockwise about the normal."""

    fmm_kernel = None

    def __init__(self, vertices, strength):
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float64)
        self.strength = np.ascontiguousarray(strength, dtype=np.float64).reshape(-1)
        n = self.vertices.shape[0]
        self.potential = np.zeros(n)
        self.gradient = np.zeros((n, 3))
        self.hessian = np.zeros((n, 9))

    @property
    def nv(self):
        return self.vertices.shape[1]

    def get_n_bodies(self):
        return self.vertices.shape[0]

    def data_per_body(self):
        return 5 + 3 * self.nv

    def strength_dims(self):
        return 1

    def has_vector_potential(self):
        return False

    def positions(self):
        return self.vertices.mean(axis=1)

    def get_position(self, i):
        return self.vertices[i].mean(axis=0)

    def source_buffer(self):
        from .workloads import panel_rows
        return panel_rows(self.vertices, self.strength)

    def set_strengths(self, strengths):
        self.strength = np.ascontiguousarray(strengths, dtype=np.float64).reshape(-1)

    def buffer_to_target_system(self, target_buffer, switch):
        n = self.get_n_bodies()
        if switch.scalar_potential:
            self.potential += target_buffer[:n, 3]
        if switch.gradient:
            self.gradient += target_buffer[:n, 4:7]
        if switch.hessian:
            self.hessian += target_buffer[:n, 7:16]

    def reset(self):
        self.potential[:] = 0.0
        self.gradient[:] = 0.0
        self.hessian[:] = 0.0


class SourcePanels(_Panels):
    """Panel{Source}: constant source density on each triangle / planar quadrilateral."""

    @property
    def fmm_kernel(self):
        return L.PANEL_SOURCE_QUAD if self.nv == 4 else L.PANEL_SOURCE_TRI


class DipolePanels(_Panels):
    """Panel{Dipole}: constant normal-doublet density on each triangle / planar quadrilateral."""

    @property
    def fmm_kernel(self):
        return L.PANEL_DIPOLE_QUAD if self.nv == 4 else L.PANEL_DIPOLE_TRI


class SourceDipolePanels(_Panels):
    """Panel{SourceDipole} (test/panels.jl:323-429): a constant source density and a constant normal-doublet
    density on each triangle / planar quadrilateral; strength (N, 2) = [source, dipole]."""

    def __init__(self, vertices, strength):
        super().__init__(vertices, np.zeros(np.asarray(vertices).shape[0]))
        self.strength = np.ascontiguousarray(strength, dtype=np.float64).reshape(-1, 2)

    @property
    def fmm_kernel(self):
        return L.PANEL_SOURCE_DIPOLE_QUAD if self.nv == 4 else L.PANEL_SOURCE_DIPOLE_TRI

    def data_per_body(self):
        return 6 + 3 * self.nv

    def strength_dims(self):
        return 2

    def source_buffer(self):
        """rows [centroid(3), radius, source strength, dipole strength, vertices]"""
        from .workloads import panel_rows
        rows = panel_rows(self.vertices, self.strength[:, 0])
        return np.ascontiguousarray(np.insert(rows, 5, self.strength[:, 1], axis=1))

