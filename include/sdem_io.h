/* sdem_io.h -- files in the reference's formats, read and written by
 * libsdem_hip.so for a context's selected box (one rank: the file's
 * particle order is ascending id).
 *
 * Restart file  RestartFor<RunName><I10.10 itime>  (Prtcl_IOAndVisu.f90:
 * 849-1031 writes it, :731-844 and :296-377 read it), native endian, no
 * record markers:
 *   int32 np, int32 nLinks
 *   np x {f64 x, y, z}
 *   np x {int32 id, int32 type, int32 usrMark}
 *   np x nreal3 x {f64 x, y, z},  nreal3 = 2 tsize + 2 rsize + 1:
 *        linVel(1..tsize), linAcc(1..tsize), theta, rotVel(1..rsize),
 *        rotAcc(1..rsize)
 *   np x int32 ncv          links of each particle (at most 20)
 *   nLinks x {int32 partner id, int32 CLid}   own records, then the
 *                                             PP records owned by a partner
 *   nTanDel x {f64 TanDelta x, y, z, f64 VelRel_Init}   record CLid (1-based)
 *
 * Visualisation dump  PartVisuFor<RunName><I10.10 itime>  (:1193-1267):
 * np x {f64 x, y, z}, then the selected fields in this order: ID (int32),
 * Diameter (f64), Type, UsrMark (int32), LinVel, LinAcc, Theta, RotVel,
 * RotAcc (first level, f64 x 3), CntctForce, Torque (f64 x 3).  Its XDMF
 * index PartVisuFor<RunName>.xmf (:64-132, :1048-1162).
 *
 * Every call returns SDEM_OK or a negative SDEM_E_* code; the context calls
 * leave their message in sdem_last_error.
 */
#ifndef SDEM_IO_H
#define SDEM_IO_H

#include <stdint.h>

#include "sdem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* PIO_Write_Restart: the state and contact records after the last step.
 * With mark_links = 1, GPrtcl_UsrMark of every particle becomes its link
 * count afterwards, as the reference's writer leaves it (:979-983; the file
 * holds the marks from before).  Writes nothing when the box is empty (as
 * the reference); SDEM_E_CAPACITY when a particle has more than 20 links
 * (the reference stops there). */
int sdem_write_restart(sdem_ctx* c, const char* path, int32_t mark_links);

/* PIO_Read_Restart + PIO_RestartCL: sets the state (radius from the
 * particle type, :809) and the contact records (CL_Add_RestartCntctlink,
 * Prtcl_CL_and_CF.f90:577-622: a link to a lower-id particle within
 * 1.5 (ri+rj) is that particle's record).  Properties, walls and fixed
 * particles must be set first. */
int sdem_read_restart(sdem_ctx* c, const char* path);

/* PIO_ReadFixedCoord (Prtcl_IOAndVisu.f90:164-291): the fixed spheres of
 * RestartDir/FixedSpheresCoord.dat, num_fix packed 36-byte records
 * {f64 x, y, z, f64 diameter, int32 type}.  Sphere k (1-based) gets the id
 * numPrtcl + k and its type's radius; the file size must be 36 num_fix
 * bytes, each diameter must match its type's within 1e-6 and each sphere
 * must lie in the domain (the reference stops otherwise).  Call after the
 * particle properties and before the state. */
int sdem_read_fixed_coord(sdem_ctx* c, const char* path, int32_t num_fix);

/* Fields of the visualisation dump (&PrtclVisuOption save_*) */
#define SDEM_VISU_ID          (1u << 0)
#define SDEM_VISU_DIAMETER    (1u << 1)
#define SDEM_VISU_TYPE        (1u << 2)
#define SDEM_VISU_USRMARK     (1u << 3)
#define SDEM_VISU_LINVEL      (1u << 4)
#define SDEM_VISU_LINACC      (1u << 5)
#define SDEM_VISU_THETA       (1u << 6)
#define SDEM_VISU_ROTVEL      (1u << 7)
#define SDEM_VISU_ROTACC      (1u << 8)
#define SDEM_VISU_CNTCTFORCE  (1u << 9)
#define SDEM_VISU_TORQUE      (1u << 10)

/* PIO_Dump_visu's binary file (forces need the per-particle force arrays
 * kept: sdem_set_tuning keep_forces = 1). */
int sdem_write_visu(sdem_ctx* c, const char* path, uint32_t fields);

/* The XDMF index, text as the reference writes it: the header with the
 * time list of ifirst-1, ifirst-1+save_visu, ... (PIO_Init_visu), one grid
 * per dump pointing into <data_base><I10.10 itime> (PIO_Write_XDMF), and
 * the tail (PIO_Final_visu). */
int sdem_xdmf_begin(const char* xmf_path, int32_t ifirst, int32_t ilast, int32_t save_visu);
int sdem_xdmf_grid(const char* xmf_path, const char* data_base, int32_t itime, int32_t np, uint32_t fields);
int sdem_xdmf_end(const char* xmf_path);

#ifdef __cplusplus
}
#endif

#endif
