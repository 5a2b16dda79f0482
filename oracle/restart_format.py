"""The reference's restart and visualisation files, restated in numpy --
TEST INFRASTRUCTURE.

Only the tests use this module, as the checker of the library's readers and
writers (sdem_amd/csrc/sdem_io.cpp, include/sdem_io.h); nothing in the
product imports it.  Every function follows the reference's Fortran
(src/DEM/Prtcl_IOAndVisu.f90, src/DEM/Prtcl_CL_and_CF.f90) for one MPI
rank, with MPI's native representation: int = int32, real = float64,
real3/real4 = 3/4 consecutive float64, little endian, no record markers.

The writer lays the file out sequentially, as PIO_Write_Restart's
Prtcl_dump calls advance `disp`; the reader finds every block by the offset
arithmetic of PIO_Read_Restart and PIO_RestartCL instead, so the two
restate the layout independently.
"""
from __future__ import annotations

import numpy as np

INT = np.dtype("<i4")
REAL = np.dtype("<f8")


def nreal3_of(tsize: int, rsize: int) -> int:
    """real3 items per particle after the position: PIO_Write_Restart
    :941-947 (2 (1 + tsize + rsize), minus 1 for GPrtcl_Pos)."""
    return 2 * (1 + tsize + rsize) - 1


# ---------------------------------------------------------------- writing
def links_of(state_ids, kind, owner, partner):
    """CL_Gather_Cntctlink_Restart (Prtcl_CL_and_CF.f90:656-685) after
    CL_Prepare_Restart (:344-359): per particle its own bucket's records
    (partner, CLid), then the counterpart list (owner of a PP record whose
    partner it is).  The bucket is a push-front list, so it is walked newest
    first: records are numbered here in input order and each bucket is
    walked from its last record back."""
    index = {int(g): i for i, g in enumerate(state_ids)}
    n = len(state_ids)
    own = [[] for _ in range(n)]
    cp = [[] for _ in range(n)]
    for r in range(len(kind)):
        clid = r + 1
        own[index[int(owner[r])]].append((int(partner[r]), clid))
        if int(kind[r]) == 1:
            cp[index[int(partner[r])]].append((int(owner[r]), clid))
    return [own[i][::-1] + cp[i][::-1] for i in range(n)]


def write_restart(path, state, contacts, tsize, rsize):
    """PIO_Write_Restart (Prtcl_IOAndVisu.f90:849-1031), one rank."""
    ids = np.asarray(state["id"])
    n = len(ids)
    links = links_of(ids, contacts["kind"], contacts["owner"], contacts["partner"])
    ncv = np.array([len(l) for l in links], INT)
    assert ncv.max(initial=0) <= 20, "so big ncv"  # CL_Count_Cntctlink :641
    parts = [np.array([n, int(ncv.sum())], INT).tobytes()]
    parts.append(np.asarray(state["posr"], REAL)[:, :3].tobytes())                      # :916-925
    im = np.stack([ids, state["ptype"], state["usr_mark"]], 1).astype(INT)               # :927-936
    parts.append(im.tobytes())
    per = [np.asarray(state["lin_vel"], REAL).reshape(n, tsize * 3),                     # :948-976
           np.asarray(state["lin_acc"], REAL).reshape(n, tsize * 3),
           np.asarray(state["theta"], REAL).reshape(n, 3),
           np.asarray(state["rot_vel"], REAL).reshape(n, rsize * 3),
           np.asarray(state["rot_acc"], REAL).reshape(n, rsize * 3)]
    parts.append(np.concatenate(per, 1).tobytes())
    parts.append(ncv.tobytes())                                                          # :978-984
    lk = [v for l in links for (g, c) in l for v in (g, c)]                             # :994-1007
    parts.append(np.array(lk, INT).tobytes())
    td = np.concatenate([np.asarray(contacts["tan_delta"], REAL).reshape(-1, 3),         # :1009-1026
                         np.asarray(contacts["vrel_init"], REAL).reshape(-1, 1)], 1)
    parts.append(td.tobytes())
    with open(path, "wb") as f:
        f.write(b"".join(parts))


# ---------------------------------------------------------------- reading
def read_restart(path, tsize, rsize, type_radius):
    """PIO_Read_Restart (Prtcl_IOAndVisu.f90:731-844) and PIO_RestartCL
    (:296-377), one rank: the state (radius from the type, :809) and, per
    particle, its links as (partner id, TanDelta xyz, VelRel_Init)."""
    raw = open(path, "rb").read()
    ib, r3b, r4b = 4, 24, 32
    np_, ncntct = np.frombuffer(raw, INT, 2, 0)
    np_, ncntct = int(np_), int(ncntct)
    disp = 2 * ib
    nreal3 = nreal3_of(tsize, rsize)
    disp_pos = disp                                   # :788-790
    disp_int = disp_pos + r3b * np_
    disp_real3 = disp_int + ib * np_ * 3
    pos = np.frombuffer(raw, REAL, 3 * np_, disp_pos).reshape(np_, 3)
    im = np.frombuffer(raw, INT, 3 * np_, disp_int).reshape(np_, 3)
    r3 = np.frombuffer(raw, REAL, 3 * nreal3 * np_, disp_real3).reshape(np_, nreal3, 3)
    k = 0
    lin_vel = r3[:, k:k + tsize]; k += tsize         # :811-816
    lin_acc = r3[:, k:k + tsize]; k += tsize
    theta = r3[:, k]; k += 1
    rot_vel = r3[:, k:k + rsize]; k += rsize
    rot_acc = r3[:, k:k + rsize]; k += rsize
    ptype = im[:, 1]
    posr = np.concatenate([pos, np.asarray(type_radius, REAL)[ptype - 1][:, None]], 1)
    state = dict(id=im[:, 0].copy(), ptype=ptype.copy(), usr_mark=im[:, 2].copy(), posr=posr,
                 lin_vel=lin_vel.copy(), lin_acc=lin_acc.copy(), theta=theta.copy(), rot_vel=rot_vel.copy(),
                 rot_acc=rot_acc.copy())
    # PIO_RestartCL's view (:329-345): pbyte uses nreal3 without the "-1"
    nreal3_cl = 2 * (1 + tsize + rsize)
    pbyte = ib * 3 + nreal3_cl * r3b
    disp = 2 * ib
    disp_ncv = disp + pbyte * np_
    disp_cl = disp_ncv + ib * np_
    disp_tan = disp_cl + 2 * ncntct * ib
    ncv = np.frombuffer(raw, INT, np_, disp_ncv)
    cl = np.frombuffer(raw, INT, 2 * ncntct, disp_cl).reshape(ncntct, 2)
    links, at = [], 0
    for i in range(np_):                               # :347-370
        mine = []
        for _ in range(int(ncv[i])):
            gid2, clid = int(cl[at, 0]), int(cl[at, 1])
            at += 1
            t = np.frombuffer(raw, REAL, 4, disp_tan + r4b * (clid - 1))
            mine.append((gid2, t[:3].copy(), float(t[3])))
        links.append(mine)
    return state, links


def add_restart_links(state, links, fixed_ids, wall_ids, periodic, lo, hi):
    """CL_Add_RestartCntctlink (Prtcl_CL_and_CF.f90:577-622) for one rank,
    with each record's kind named: a link to a lower-id particle within
    1.5 (ri + rj) is skipped (that particle holds the record); a link to a
    moving particle across a periodic boundary is the particle's own PPG
    record; to a higher-id moving particle otherwise a PP record; fixed
    particles give PPFix, wall ids PW.  Returns the records as
    (kind, owner, partner, tan_delta, vrel_init) arrays."""
    ids = np.asarray(state["id"])
    posr = np.asarray(state["posr"])
    index = {int(g): i for i, g in enumerate(ids)}
    fixed, walls = {int(g) for g in fixed_ids}, {int(g) for g in wall_ids}
    L = np.asarray(hi, float) - np.asarray(lo, float)
    rec = []
    for i, mine in enumerate(links):
        g1 = int(ids[i])
        for g2, td, vri in mine:
            if g2 in index:
                j = index[g2]
                d = posr[i, :3] - posr[j, :3]
                dist = np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) - 1.5 * (posr[i, 3] + posr[j, 3])
                if g1 > g2 and dist <= 0.0:
                    continue
                across = any(periodic[k] and abs(d[k]) > 0.5 * L[k] for k in range(3))
                if across:
                    kind = 2
                elif g1 < g2:
                    kind = 1
                else:
                    continue
            elif g2 in fixed:
                kind = 3
            elif g2 in walls:
                kind = 4
            else:
                raise ValueError(f"link to unknown id {g2}")
            rec.append((kind, g1, g2, td, vri))
    if not rec:
        return dict(kind=np.zeros(0, np.int32), owner=np.zeros(0, np.int32), partner=np.zeros(0, np.int32),
                    tan_delta=np.zeros((0, 3)), vrel_init=np.zeros(0))
    return dict(kind=np.array([r[0] for r in rec], np.int32), owner=np.array([r[1] for r in rec], np.int32),
                partner=np.array([r[2] for r in rec], np.int32),
                tan_delta=np.array([r[3] for r in rec]), vrel_init=np.array([r[4] for r in rec]))


# ------------------------------------------------------------ visu dumps
VISU_ORDER = ("ID", "Diameter", "Type", "UsrMark", "LinVel", "LinAcc", "Theta", "RotVel", "RotAcc",
              "CntctForce", "Torque")
# (XDMF attribute type, data type, dims, precision): PIO_Write_XDMF :1086-1129
VISU_META = {"ID": ("Scalar", "Int", 1, 4), "Diameter": ("Scalar", "Float", 1, 8),
             "Type": ("Scalar", "Int", 1, 4), "UsrMark": ("Scalar", "Int", 1, 4),
             "LinVel": ("Vector", "Float", 3, 8), "LinAcc": ("Vector", "Float", 3, 8),
             "Theta": ("Vector", "Float", 3, 8), "RotVel": ("Vector", "Float", 3, 8),
             "RotAcc": ("Vector", "Float", 3, 8), "CntctForce": ("Vector", "Float", 3, 8),
             "Torque": ("Vector", "Float", 3, 8)}


def read_visu(path, np_, fields):
    """PIO_Dump_visu's file (Prtcl_IOAndVisu.f90:1193-1267): positions, then
    each selected field in VISU_ORDER."""
    raw = open(path, "rb").read()
    out, disp = {}, 0
    out["Pos"] = np.frombuffer(raw, REAL, 3 * np_, disp).reshape(np_, 3)
    disp += 24 * np_
    for name in VISU_ORDER:
        if name not in fields:
            continue
        _, dt, dims, prec = VISU_META[name]
        a = np.frombuffer(raw, INT if dt == "Int" else REAL, dims * np_, disp)
        out[name] = a.reshape(np_, dims) if dims == 3 else a
        disp += dims * prec * np_
    assert disp == len(raw), (disp, len(raw))
    return out


def _I(v, w, zero=False):
    """Fortran Iw / Iw.w output (asterisks when the value does not fit)."""
    s = f"{v:0{w}d}" if zero else f"{v:{w}d}"
    return "*" * w if len(s) > w else s


def xdmf_begin(ifirst, ilast, save_visu):
    """PIO_Init_visu's header (Prtcl_IOAndVisu.f90:89-112)."""
    nflds = (ilast - ifirst + 1) // save_visu + 1
    t = "".join(_I((f - 1) * save_visu + ifirst - 1, 9) for f in range(1, nflds + 1))
    return ('<?xml version="1.0" ?>\n'
            '<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>\n'
            '<Xdmf xmlns:xi="http://www.w3.org/2001/XInclude" Version="2.0">\n'
            '<Domain>\n'
            + " " * 4 + '<Grid Name="TimeSeries" GridType="Collection" CollectionType="Temporal">\n'
            + " " * 8 + '<Time TimeType="List">\n'
            + " " * 12 + f'<DataItem Format="XML" NumberType="Int" Dimensions="{_I(nflds, 6)}">\n'
            + " " * 12 + t + " " * 12 + "</DataItem>\n"
            + " " * 8 + "</Time>\n")


def xdmf_grid(data_base, itime, np_, fields):
    """PIO_Write_XDMF + Write_XDMF_One (Prtcl_IOAndVisu.f90:1048-1162)."""
    fname = data_base + _I(itime, 10, True)
    s = " " * 8 + f'<Grid Name="T{_I(itime, 10, True)}" GridType="Uniform">\n'
    s += " " * 12 + f'<Topology TopologyType="Polyvertex" NodesPerElement="{_I(np_, 9)}"/>\n'
    s += " " * 12 + '<Geometry GeometryType="XYZ">\n'
    disp = 0
    s += (" " * 16 + '<DataItem Format="Binary" DataType="Float" Precision="8" Endian="Native" Dimensions="'
          + _I(3, 2) + _I(np_, 9) + '" Seek="' + _I(disp, 15) + '">\n')
    disp += np_ * 3 * 8
    s += " " * 20 + fname + "\n" + " " * 16 + "</DataItem>\n" + " " * 12 + "</Geometry>\n"
    for name in VISU_ORDER:
        if name not in fields:
            continue
        attr, dt, dims, prec = VISU_META[name]
        s += " " * 12 + f'<Attribute Type="{attr}" Center="Node" Name="{name}">\n'
        s += (" " * 16 + f'<DataItem Format="Binary" DataType="{dt}" Precision="{prec}" Endian="Native" '
              f'Dimensions="{_I(dims, 2)}{_I(np_, 9)}" Seek="{_I(disp, 15)}">\n')
        disp += np_ * dims * prec
        s += " " * 20 + fname + "\n" + " " * 16 + "</DataItem>\n" + " " * 12 + "</Attribute>\n"
    return s + "        </Grid>\n"


def xdmf_end():
    """PIO_Final_visu's tail (Prtcl_IOAndVisu.f90:1183-1185)."""
    return "    </Grid>\n</Domain>\n</Xdmf>\n"
