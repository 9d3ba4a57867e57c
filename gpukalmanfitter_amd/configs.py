"""The five benchmark / parity workloads of the fitter.

Units: mm, GeV, mm/c.  Each entry is the generator configuration of
``gpukalmanfitter_amd.generate`` plus the field description.

1  Plain FP64 fit, CPU-runnable: 10,000 tracks, 10 x-normal silicon planes at
   x = 30..300 mm (1 m square), uniform Bz = 2 T, no material.
2  The same detector and generator with 1,048,576 tracks (FP64 and
   FP32-mixed fits).
3  Ragged batch: 2,000,000 tracks, U{5..30} planes at 25 mm pitch, 5 % holes,
   101x101x201 FP32 grid field over a 2 x 2 x 4 m box,
   Bz = 2 T (1 - 0.1 (r/1 m)^2), Br = 0.05 T (z / 2 m).
4  Alignment: 262,144 tracks on the config-1 detector simulated with seeded
   misalignments (shifts 20 um, rotations 0.1 mrad) and a field scale
   1 +- 0.01.
5  Material: 4,194,304 tracks, U{8..40} planes of 300 um silicon at 25 mm
   pitch, p in [0.5, 20] GeV, config-3 field.
"""
import math

TESLA = 0.000299792458
NS = 1e-9 * 299792458000.0
MUON_MASS = 105.6583755 * 1e-3

_BASE = dict(
    n_tracks=10000,
    seed=1234567890,
    n_planes=10,
    min_planes=10,
    max_planes=10,
    material=0,
    first_x=30.0,
    pitch=30.0,
    half_u=500.0,
    half_v=500.0,
    thickness=0.0,
    hole_prob=0.0,
    p_min=1.0,
    p_max=10.0,
    phi_min=-0.3,
    phi_max=0.3,
    eta_min=-0.3,
    eta_max=0.3,
    vertex_sigma=0.010,
    hit_sigma_u=0.050,
    hit_sigma_v=0.050,
    sigma_d0=0.100,
    sigma_z0=0.100,
    sigma_phi=1e-3,
    sigma_theta=1e-3,
    sigma_p_rel=0.05,
    sigma_t=5.0 * NS,
    mass=MUON_MASS,
    align_shift_sigma=0.0,
    align_rot_sigma=0.0,
    true_field_scale=1.0,
)

# kfg_config.material: the reference's simulator, or the fit's own material
# model (Gaussian projected scattering and Gaussian mean-Bethe loss, no
# Bethe-Heitler) for pull tests of the material update, or that model's energy
# loss alone (for fits without multiple scattering)
MATERIAL_REFERENCE = 1
MATERIAL_FIT_MODEL = 2
MATERIAL_FIT_MODEL_ELOSS = 3

GRID_FIELD = dict(
    kind="grid",
    n=(101, 101, 201),
    lo=(-1000.0, -1000.0, -2000.0),
    hi=(1000.0, 1000.0, 2000.0),
    b0=2.0 * TESLA,
    br=0.05 * TESLA,
)
CONST_FIELD = dict(kind="const", b=(0.0, 0.0, 2.0 * TESLA))


def _cfg(**kw):
    c = dict(_BASE)
    c.update(kw)
    return c


CONFIGS = {
    1: dict(gen=_cfg(n_tracks=10000, seed=1234567891), field=CONST_FIELD,
            name="config1-fp64-10k"),
    2: dict(gen=_cfg(n_tracks=1048576, seed=1234567892), field=CONST_FIELD,
            name="config2-1M"),
    3: dict(gen=_cfg(n_tracks=2000000, seed=1234567893, n_planes=30, min_planes=5,
                     max_planes=30, first_x=25.0, pitch=25.0, hole_prob=0.05),
            field=GRID_FIELD, name="config3-ragged-grid"),
    4: dict(gen=_cfg(n_tracks=262144, seed=1234567894, align_shift_sigma=0.020,
                     align_rot_sigma=1e-4), field=CONST_FIELD,
            field_scale_sigma=0.01, name="config4-alignment"),
    5: dict(gen=_cfg(n_tracks=4194304, seed=1234567895, n_planes=40, min_planes=8,
                     max_planes=40, first_x=25.0, pitch=25.0, material=MATERIAL_REFERENCE,
                     thickness=0.300, p_min=0.5, p_max=20.0),
            field=GRID_FIELD, name="config5-material"),
}


# Fit options of the workloads (kf_options fields).  The reference's
# overstep limit of 0.01 mm (EigenStepper.hpp:370) loses every track whose
# curvature carries it more than 0.01 mm past the straight-line plane estimate
# (19 % of config-1 tracks: mu+- at |phi| <= 0.3 rad bending towards the plane
# normal); the workloads use 1 mm so every track is fitted.  The default of
# kf_options keeps the reference's 0.01 mm.
WORKLOAD_FIT_OPTIONS = dict(overstep_limit=1.0)


def theta_from_eta(eta):
    return 2.0 * math.atan(math.exp(-eta))
