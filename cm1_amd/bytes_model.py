"""Algorithmic HBM bytes of the CM1 large step (frozen table).

Counting rule: 4 B x (interior cells) for every distinct full-size 3-D
field a routine must read, plus every field it must write, each counted
once per invocation.  Routine-internal scratch (advection fluxes, fpk,
Thomas coefficients, dum*) and 1-D/2-D metrics are excluded.  Where
kernels are fused across routine boundaries the fused kernel's distinct
field count is used.

Per-invocation field counts of the reference routines (read + write),
from the loops they come from:

==========================  ======  =========================================
routine                     fields  reference loops
==========================  ======  =========================================
advu / advv / advw             7    src/adv.F:911-3610 (vel, rru, rrv, rrw,
                                    divx, ten in, ten out)
advs                           7    src/adv.F:14-903 (8 with pdef: + s0)
sound substep                 27    src/sound.F:307-835 (u,v,w,ppd,pp3d,
                                    rru,rrv,rrw R+W; uten,vten,wten,thv,
                                    ppten,ppterm,piadv,pk1,pk2,qk,mm R)
sound setup                    7    src/sound.F:117-180
sound prep                  6 / 13  src/sound.F:191-250 (nrk=1 / nrk>=2)
==========================  ======  =========================================

Per large step: 534 fields per cell outside the acoustic substeps plus 27
per substep (2+3+6 = 11 substeps at nsound=6, 3+4+8 = 15 at nsound=8,
src/sound.F:84-104).
"""

BYTES_PER_FIELD = 4

# fields per invocation
ADV_MOMENTUM = 7
ADVS = 7
ADVS_PDEF = 8
SOUND_SUBSTEP = 27
SOUND_SETUP = 7
SOUND_PREP = {1: 6, 2: 13, 3: 13}
STEP_BASE = 534

# the fused acoustic substep kernel (sound_sub) evaluates pk1, pk2, qk and
# mm on the fly, so it touches 23 distinct fields: u, v, w, ppd, pp3d, rru,
# rrv, rrw read and written (16) and uten, vten, wten, thv, ppten, ppterm,
# piadv read (7).  The roofline's algorithmic bytes of one launch are the
# routine's 27 fields (SOUND_SUBSTEP); the 23-field figure is reported
# beside it.
SOUND_SUBSTEP_FUSED = 23


def substeps_per_step(nsound):
    """Acoustic substeps in one RK3 large step (src/sound.F:84-104)."""
    n1 = int(round(nsound / 3.0))
    # the first stage rounds nsound/3 and adds one substep if the result
    # would exceed the acoustic time-step limit dt/nsound
    if 1.0 / (n1 * 3.0) > 1.0 / nsound:
        n1 += 1
    return n1 + int(0.5 * nsound) + nsound


def fields_per_step(nsound):
    return STEP_BASE + SOUND_SUBSTEP * substeps_per_step(nsound)


def bytes_per_cell_step(nsound):
    return BYTES_PER_FIELD * fields_per_step(nsound)
