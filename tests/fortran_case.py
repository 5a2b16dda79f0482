"""A small case in the reference's .prm format, for the Fortran host tests:
the text of the file and the same case as a scenario for the Python side."""
import numpy as np

from sdem_amd import native, scenarios

CASE = dict(num=1500, steps=40, dt=2.0e-5, box=(0.04, 0.06, 0.04), periodic=(True, False, True),
            diam=(2.0e-3, 3.0e-3), dens=(2500.0, 2600.0), E=5.0e7, nu=0.3, bins=(2.0, 1.0))


def prm_text(c=CASE, cf_type=2, pi=2, pri=2, restart=False, ifirst=1, ilast=None, save_visu=100000,
             backup=100000, visu=None) -> str:
    """visu: None (no &PrtclVisuOption) or a dict of its logicals."""
    T = lambda b: "T" if b else "F"  # noqa: E731
    bx = c["box"]
    ilast = c["steps"] if ilast is None else ilast
    visu_nml = ""
    if visu is not None:
        visu_nml = "&PrtclVisuOption\n" + "".join(f"  {k} = {T(v)}\n" for k, v in visu.items()) + "/\n"
    return f"""!
! settling of two sphere sizes between two infinite planes (y = 0, y = top)
!
&DEMOptions
  RestartFlag = {T(restart)}
  numPrtclFix = 0
  numPrtcl = {c['num']}
  dtDEM = {c['dt']:.6E}
  ifirstDEM = {ifirst}
  ilastDEM = {ilast}
  gravity = 0.000 -9.81 0.0
  minpoint = 0.000 0.000 0.000
  maxpoint = {bx[0]} {bx[1]} {bx[2]}
  isPeriodic = {T(c['periodic'][0])},{T(c['periodic'][1])},{T(c['periodic'][2])}
  Prtcl_cs_ratio = 1.0
  CS_Method = 1
  CF_Type = {cf_type}
  CT_Model = 1
  PI_Method = {pi}
  PRI_Method = {pri}
  numPrtcl_Type = 2
  numWall_type = 1
  CntctList_Size = 6
  CS_numlvls = 0
  Base_wall_id = 10000000
  Wall_max_update_iter = 100
  Wall_neighbor_ratio = 2.00
  RunName = "Settling"
  Res_Dir = "./"
  RestartDir = "./"
  SaveVisuDEM = {save_visu}
  BackupFreqDEM = {backup}
  Cmd_LFile_Freq = 10
  LF_file_lvl = 5
  LF_cmdw_lvl = 3
  GeometrySource = 1
  Geom_Dir = "DEMGeom.stl"
/

&ParticlePhysicalProperty
  Bin_Divided = {c['bins'][0]} {c['bins'][1]}
  Density = {c['dens'][0]} {c['dens'][1]}
  Diameter = {c['diam'][0]:.4E} {c['diam'][1]:.4E}
  YoungsModulus_P = {c['E']:.4E} {c['E']:.4E}
  PoissonRatio_P = {c['nu']} {c['nu']}
  FrictionCoe_s_PP = 0.40
  FrictionCoe_k_PP = 0.35
  FrictionCoe_Roll_PP = 0.05
  RestitutionCoe_n_PP = 0.60
  RestitutionCoe_t_PP = 0.80
/

&LinearContactModel
  kn = 2.0E+4
  kt_over_kn = 0.8
/

&WallPhysicalProperty
  YoungsModulus_W = {c['E']:.4E}
  PoissonRatio_W = {c['nu']}
  FrictionCoe_s_PW = 0.45
  FrictionCoe_k_PW = 0.40
  FrictionCoe_Roll_PW = 0.05
  RestitutionCoe_n_PW = 0.50
  RestitutionCoe_t_PW = 0.80
/

&ParticleMakingOption
  MkPrtclMinpoint = 0.000 0.000 0.000
  MkPrtclMaxpoint = {bx[0]} {bx[1]} {bx[2]}
  Fill_order = "x","z","y"
  Distance_Ratio = 1.05
  VelMag = 0.3
/

&GeometryMakingNumPlane
  nplane = 2
/

&GeometryMakingParam
  user_id = 1 2
  wall_Type = 1 1
  bothSide = F F
  IsInfinite = T T
  Point1 = 0.00 0.00 0.00, 0.00 {bx[1]} 0.00,
  Point2 = 0.00 0.00 {bx[2]}, {bx[0]} {bx[1]} 0.00,
  Point3 = {bx[0]} 0.00 {bx[2]}, {bx[0]} {bx[1]} {bx[2]},
  Point4 = {bx[0]} 0.00 0.00, 0.00 {bx[1]} {bx[2]},
  TraVel = 0.00 0.00 0.00, 0.00 0.00 0.00,
/
""" + visu_nml


def read_dump(path: str) -> dict:
    """The Fortran driver's raw state dump (see Prtcl_DEMSystem.dump_state)."""
    raw = open(path, "rb").read()
    n, ts, rs = np.frombuffer(raw, np.int32, 3)
    off = 12
    ints = np.frombuffer(raw, np.int32, 3 * n, off).reshape(3, n)
    off += 12 * n
    d = np.frombuffer(raw, np.float64, offset=off)
    k = 0

    def take(m, shape):
        nonlocal k
        a = d[k:k + m].reshape(shape)
        k += m
        return a.copy()
    posr = take(4 * n, (n, 4))
    lv = take(3 * ts * n, (n, ts, 3))
    la = take(3 * ts * n, (n, ts, 3))
    th = take(3 * n, (n, 3))
    rv = take(3 * rs * n, (n, rs, 3))
    ra = take(3 * rs * n, (n, rs, 3))
    return dict(id=ints[0].copy(), ptype=ints[1].copy(), usr_mark=ints[2].copy(), posr=posr, lin_vel=lv,
                lin_acc=la, theta=th, rot_vel=rv, rot_acc=ra)


def scenario(state: dict, c=CASE, cf_type=2, pi=2, pri=2) -> dict:
    """The same case for native.build_engine, from a dumped initial state."""
    bx = c["box"]
    opts = scenarios.base_options(dt=c["dt"], gravity=(0.0, -9.81, 0.0), domain_max=tuple(bx),
                                  is_periodic=tuple(int(p) for p in c["periodic"]), num_prtcl=c["num"],
                                  cf_type=cf_type, pi_method=pi, pri_method=pri)
    walls = [dict(p1=(0, 0, 0), p2=(0, 0, bx[2]), p3=(bx[0], 0, bx[2]), p4=(bx[0], 0, 0), user_id=1,
                  wall_type=1, both_side=False, is_infinite=True, trans_vel=(0.0, 0.0, 0.0)),
             dict(p1=(0, bx[1], 0), p2=(bx[0], bx[1], 0), p3=(bx[0], bx[1], bx[2]), p4=(0, bx[1], bx[2]),
                  user_id=2, wall_type=1, both_side=False, is_infinite=True, trans_vel=(0.0, 0.0, 0.0))]
    sc = dict(options=opts,
              ptypes=dict(diameter=list(c["diam"]), density=list(c["dens"]), youngs=[c["E"]] * 2,
                          poisson=[c["nu"]] * 2),
              pp=(0.40, 0.35, 0.05, 0.60, 0.80),
              wtypes=dict(youngs=[c["E"]], poisson=[c["nu"]]), pw=(0.45, 0.40, 0.05, 0.50, 0.80),
              walls=walls, state=state)
    if cf_type == native.CFT_LINEAR:
        sc["linear"] = (2.0e4, 0.8)
    return sc


# ---------------------------------------------------------------------------
# an hourglass case in the format of the reference's Input/SandBox_DEM.prm:
# the program's own geometry (GeometrySource = 0) and the sand box's
# materials; 3 mm spheres filled into the neck from its walls (the first
# lattice centres 1.05 r from the fill box's sides) and from just above its
# lower edge, moving at 0.3 m/s
# ---------------------------------------------------------------------------
SANDBOX = dict(num=1200, steps=300, dt=1.0e-5, diam=3.0e-3, dens=2500.0, E=1.0e6, nu=0.23,
               mk_min=(-0.015, -0.0515, -0.015), mk_max=(0.015, 0.12, 0.015))


def sandbox_prm_text(c=SANDBOX) -> str:
    mn, mx = c["mk_min"], c["mk_max"]
    return f"""!
! hourglass: the program's own walls (GeometrySource = 0), spheres in the neck
!
&DEMOptions
  RestartFlag = F
  numPrtclFix = 0
  numPrtcl = {c['num']}
  dtDEM = {c['dt']:.6E}
  ifirstDEM = 1
  ilastDEM = {c['steps']}
  gravity = 0.0 -9.81 0.0
  minpoint = -0.075 -0.30 -0.075
  maxpoint = 0.075 0.30 0.075
  isPeriodic = F,F,F
  Prtcl_cs_ratio = 1.0
  CS_Method = 1
  CF_Type = 2
  CT_Model = 1
  PI_Method = 3
  PRI_Method = 3
  numPrtcl_Type = 1
  numWall_type = 1
  CntctList_Size = 6
  CS_numlvls = 0
  Base_wall_id = 10000000
  Wall_max_update_iter = 100
  Wall_neighbor_ratio = 2.00
  RunName = "SandBox"
  Res_Dir = "./"
  RestartDir = "./"
  SaveVisuDEM = 100000
  BackupFreqDEM = 100000
  Cmd_LFile_Freq = 50
  LF_file_lvl = 5
  LF_cmdw_lvl = 3
  GeometrySource = 0
  Geom_Dir = "DEMGeom.stl"
/

&ParticlePhysicalProperty
  Bin_Divided = 1
  Density = {c['dens']}
  Diameter = {c['diam']}
  YoungsModulus_P = {c['E']:.4E}
  PoissonRatio_P = {c['nu']}
  FrictionCoe_s_PP = 0.40
  FrictionCoe_k_PP = 0.40
  FrictionCoe_Roll_PP = 0.10
  RestitutionCoe_n_PP = 0.30
  RestitutionCoe_t_PP = 0.80
/

&WallPhysicalProperty
  YoungsModulus_W = {c['E']:.4E}
  PoissonRatio_W = {c['nu']}
  FrictionCoe_s_PW = 0.40
  FrictionCoe_k_PW = 0.40
  FrictionCoe_Roll_PW = 0.10
  RestitutionCoe_n_PW = 0.30
  RestitutionCoe_t_PW = 0.80
/

&ParticleMakingOption
  MkPrtclMinpoint = {mn[0]} {mn[1]} {mn[2]}
  MkPrtclMaxpoint = {mx[0]} {mx[1]} {mx[2]}
  Fill_order = "x","z","y"
  Distance_Ratio = 1.05
  VelMag = 0.3
/
"""


def hourglass_walls():
    """The 22 walls of GeometrySource = 0 from the fixture written out of the
    reference's source (tests/golden/make_hourglass_fixture.py)."""
    import json
    import os

    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "hourglass_quads.json")) as f:
        quads = json.load(f)["quads"]
    return [dict(p1=tuple(q["corners"][0]), p2=tuple(q["corners"][1]), p3=tuple(q["corners"][2]),
                 p4=tuple(q["corners"][3]), user_id=q["user_id"], wall_type=q["wall_type"],
                 both_side=q["both_side"], is_infinite=q["infinite"], trans_vel=(0.0, 0.0, 0.0))
            for q in quads]


def sandbox_scenario(state: dict, c=SANDBOX) -> dict:
    """The hourglass case for native.build_engine, from a dumped state."""
    opts = scenarios.base_options(dt=c["dt"], gravity=(0.0, -9.81, 0.0), domain_min=(-0.075, -0.30, -0.075),
                                  domain_max=(0.075, 0.30, 0.075), num_prtcl=c["num"], cf_type=native.CFT_NLIN,
                                  pi_method=native.PIM_AB3, pri_method=native.PIM_AB3)
    return dict(options=opts,
                ptypes=dict(diameter=[c["diam"]], density=[c["dens"]], youngs=[c["E"]], poisson=[c["nu"]]),
                pp=(0.40, 0.40, 0.10, 0.30, 0.80),
                wtypes=dict(youngs=[c["E"]], poisson=[c["nu"]]), pw=(0.40, 0.40, 0.10, 0.30, 0.80),
                walls=hourglass_walls(), state=state)
