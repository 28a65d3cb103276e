"""Initial conditions of the workloads (Sod, Sedov-Taylor, Kelvin-Helmholtz)
and randomised states used by the tests, on flat lattice meshes."""
import numpy as np

from rich_amd import mesh as rmesh

GAMMA = 5.0 / 3.0


def sod_mesh(nx=128, ny=16, seed=7, flip_seed=0):
    return rmesh.lattice(nx, ny, (0.0, 1.0, 0.0, 0.125), jitter=1e-3, seed=seed, bc=rmesh.RIGID,
                         flip_seed=flip_seed)


def sod_cells(m):
    x = m.pt_x[:m.n_cells]
    c = np.zeros((4, m.n_cells))
    c[0] = np.where(x < 0.5, 1.0, 0.125)
    c[1] = np.where(x < 0.5, 1.0, 0.1)
    return c


def sedov_mesh(n=64, seed=11, flip_seed=0, bc=rmesh.RIGID):
    return rmesh.lattice(n, n, (-1.0, 1.0, -1.0, 1.0), jitter=1e-3, seed=seed, bc=bc,
                         flip_seed=flip_seed)


def sedov_cells(m, energy=1.0, radius=0.02, gamma=GAMMA, rho=1.0, p_amb=1e-5):
    """Ambient (rho, p_amb); `energy` deposited as uniform thermal energy
    density over the cells whose generating point has r < radius (at least
    the one nearest the origin)."""
    N = m.n_cells
    r = np.hypot(m.pt_x[:N], m.pt_y[:N])
    hot = r < radius
    if not hot.any():
        hot = r <= r.min()
    c = np.zeros((4, N))
    c[0] = rho
    c[1] = p_amb
    c[1, hot] = (gamma - 1.0) * energy / m.volume[hot].sum()
    return c


def kh_mesh(nx=2048, ny=1024, seed=1, flip_seed=0):
    return rmesh.lattice(nx, ny, (0.0, 2.0, 0.0, 1.0), jitter=1e-3, seed=seed,
                         bc=rmesh.PERIODIC, flip_seed=flip_seed)


def kh_cells(m):
    """Band |y-0.5|<0.25: rho=2, vx=+0.5; outside rho=1, vx=-0.5; p=2.5;
    vy = 0.01 sin(4 pi x) exp(-(y-0.25)^2/0.005 - (y-0.75)^2/0.005)."""
    N = m.n_cells
    x, y = m.pt_x[:N], m.pt_y[:N]
    band = np.abs(y - 0.5) < 0.25
    c = np.zeros((4, N))
    c[0] = np.where(band, 2.0, 1.0)
    c[1] = 2.5
    c[2] = np.where(band, 0.5, -0.5)
    c[3] = 0.01 * np.sin(4 * np.pi * x) * np.exp(-(y - 0.25) ** 2 / 0.005 - (y - 0.75) ** 2 / 0.005)
    return c


def rough_cells(m, seed=0):
    """Smooth waves + a diagonal discontinuity + noise: exercises the
    unlimited, limited and shocked branches of the slope limiter."""
    rng = np.random.default_rng(seed)
    N = m.n_cells
    x, y = m.pt_x[:N], m.pt_y[:N]
    x0, x1 = m.pt_x[:N].min(), m.pt_x[:N].max()
    y0, y1 = m.pt_y[:N].min(), m.pt_y[:N].max()
    u = (x - x0) / (x1 - x0 + 1e-300)
    v = (y - y0) / (y1 - y0 + 1e-300)
    jump = (u + 0.7 * v) > 0.8
    c = np.zeros((4, N))
    c[0] = 1.0 + 0.3 * np.sin(6.0 * u) * np.cos(5.0 * v) + 1.5 * jump + 0.05 * rng.random(N)
    c[1] = 1.0 + 0.2 * np.cos(7.0 * u + 2.0 * v) + 3.0 * jump + 0.05 * rng.random(N)
    c[2] = 0.4 * np.sin(4.0 * v) - 0.6 * jump + 0.1 * rng.standard_normal(N)
    c[3] = 0.3 * np.cos(3.0 * u) + 0.5 * jump + 0.1 * rng.standard_normal(N)
    return c


def small_meshes():
    """(name, mesh) pairs covering both ghost generators and both face
    orientations."""
    return [
        ("rigid", rmesh.lattice(24, 20, (0.0, 1.2, 0.0, 1.0), jitter=1e-3, seed=3, bc=rmesh.RIGID)),
        ("rigid_flip", rmesh.lattice(20, 16, (-1.0, 1.0, -0.8, 0.8), jitter=2e-2, seed=4,
                                     bc=rmesh.RIGID, flip_seed=5)),
        ("periodic", rmesh.lattice(32, 16, (0.0, 2.0, 0.0, 1.0), jitter=1e-3, seed=5,
                                   bc=rmesh.PERIODIC)),
        ("periodic_flip", rmesh.lattice(16, 24, (0.0, 1.0, 0.0, 1.5), jitter=2e-2, seed=6,
                                        bc=rmesh.PERIODIC, flip_seed=7)),
    ]
