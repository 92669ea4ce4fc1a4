/* rslots.h - meaning of the parameter slots of ramp_part_desc per component
 * type (p[] floats, ni[] integers, tab_off[] tables).  Parsed by the Python
 * package (ramp_amd/_slots.py) so host descriptors and kernels agree.
 * Values are float32 conversions of the reference wrappers' arguments unless
 * noted; the citation gives the reference kernel argument list.
 */
#ifndef RAMP_RSLOTS_H
#define RAMP_RSLOTS_H

/* sources: flat_mod.cl:4-8, cont_mod.cl:11-16, isis_mod.cl:15-21 */
#define SP_MOD_W 0
#define SP_MOD_H 1
#define SP_TGT_W 2
#define SP_TGT_H 3
#define SP_TGT_D 4
#define SP_EMIN 5
#define SP_EMAX 6
#define SP_T1 7
#define SP_I1 8
#define SP_T2 9
#define SP_I2 10
#define SP_T3 11
#define SP_I3 12
#define SP_STR_AREA 13
#define SP_TAIL_JOIN 14       /* MCont extension: 1/E tail above TAIL_JOIN*kT1 (0 = off) */
#define SP_ISIS_TOTAL 7
#define SP_ISIS_STR_AREA 8
#define SP_ISIS_TOFFSET 9
#define SI_ISIS_NT 0
#define SI_ISIS_NE 1
#define ST_ISIS_FLUX 0
#define ST_ISIS_TBINS 1
#define ST_ISIS_EBINS 2
#define ST_ISIS_EINT 3
#define SP_LMIN 5             /* MLambda / MParallel: wavelength range (Angstrom) */
#define SP_LMAX 6
#define SP_LNORM 7            /* MLambda: (lmax-lmin)/ln(lmax/lmin) = 1/E[1/lambda] */

/* geometry */
#define GP_W 0                /* plane.cl:5-6, box.cl:5-6 */
#define GP_H 1
#define GP_D 2
#define GI_ORIENT 0
#define GP_R 0                /* sphere.cl:5, cylinder_ext.cl:5, banana.cl:5-6 */
#define GP_HEIGHT 1
#define GP_MINTH 2
#define GP_MAXTH 3
#define GP_XMIN 0             /* mesh.cl:5-8 */
#define GP_XMAX 1
#define GP_YMIN 2
#define GP_YMAX 3
#define GP_ZMIN 4
#define GP_ZMAX 5
#define GI_NTRI 0
#define GI_INTERIOR 1
#define GI_NNODES 2           /* derived: nodes of one BVH layout (filled by libramp) */
#define GT_POINTS 0
#define GT_BVH 1              /* device-side BVH (filled by libramp, unused by the oracle) */

/* scatter: chopper.cl:5-8 */
#define CP_SLIT_W 0
#define CP_RADIUS 1
#define CP_FREQ 2
#define CP_PHASE 3
#define CP_JITTER 4
#define CI_NSLITS 0
/* guide.cl:10-16 */
#define GU_W1 0
#define GU_H1 1
#define GU_W2 2
#define GU_H2 3
#define GU_L 4
#define GU_R0 5
#define GU_QC 6
#define GU_ALPHA 7
#define GU_M 8
#define GU_W 9
#define GUI_MAXB 0
/* isotropic.cl:75-112: channel c in {0 coh, 1 inc, 2 mag} */
#define IS_RHO(c) (3 * (c) + 0)
#define IS_SABS(c) (3 * (c) + 1)
#define IS_SSCAT(c) (3 * (c) + 2)
#define IS_TEMP 9
#define ISI_NQ(c) (2 * (c) + 0)
#define ISI_NW(c) (2 * (c) + 1)
#define ISI_TRANSMIT 6
/* powder1.cl:6-8 */
#define P1_SSCAT_V2 0
#define P1_SABS_V 1
#define P1_Q 2
#define P1_DPHI 3
/* powderN.cl:6-9 */
#define PN_SABS_V 0
#define PN_DPHI 1
#define PNI_NLINES 0
#define PNI_TRANSMIT 1
#define PNT_SSCAT_V2 0
#define PNT_Q 1
/* monochromator.cl:22-29 */
#define MO_SLAB_W 0
#define MO_SLAB_H 1
#define MO_GAP 2
#define MO_MOS_H 3
#define MO_MOS_V 4
#define MO_R0 5
#define MO_D 6
#define MO_RAD_V 7
#define MO_RAD_H 8
#define MOI_NH 0
#define MOI_NV 1
#define MOI_GLEN 2
#define MOT_GX 0
#define MOT_GW 1
/* detector1D.cl:20-24 */
#define D1_MIN 0
#define D1_STEP 1
#define D1_MAX 2
#define D1I_VAR 0
#define D1I_NBINS 1
#define D1I_RESTORE 2
/* detector2D.cl:31-37 */
#define D2_MIN1 0
#define D2_STEP1 1
#define D2_MAX1 2
#define D2_MIN2 3
#define D2_STEP2 4
#define D2_MAX2 5
#define D2I_N1 0
#define D2I_N2 1
#define D2I_VAR1 2
#define D2I_VAR2 3
#define D2I_RESTORE 4
/* SIncoherent (no reference counterpart) */
#define IN_SIGMA_S 0          /* 1/cm */
#define IN_SIGMA_A 1          /* 1/cm at IN_LREF */
#define IN_LREF 2             /* Angstrom */
#define IN_FOCUS_R 3          /* > 0: scatter into the cone towards the target */
#define IN_U 4                /* 4..6 unit vector towards the target */
#define IN_E1 7               /* 7..9 unit vector perpendicular to it */
#define IN_E2 10              /* 10..12 u x e1 */
#define IN_OMC 13             /* 1 - cos(cone half-angle) */
#define IN_WF 14              /* weight factor: solid-angle fraction (1 - cos)/2 */
/* SPowderDiff (no reference counterpart): model parameters (differentiated)
 * and the proposal they are sampled from (fixed) */
#define PD_SINC 0             /* incoherent Sigma_s, 1/cm */
#define PD_SABS 1             /* absorption Sigma_a at PD_LREF, 1/cm */
#define PD_LREF 2             /* Angstrom */
#define PD_U 3                /* Debye-Waller U, A^2 */
#define PD_STRAIN 4           /* relative rms width of a line's d-spacing */
#define PD_SINC0 5            /* proposal Sigma_s */
#define PD_SABS0 6            /* proposal Sigma_a */
#define PDI_NLINES 0
#define PDT_LINES 0           /* PD_LW floats per line: */
#define PD_LW 8
#define PDL_SSV2 0            /*   model sigma * v^2 (powderN convention, with exp(-U q^2)) */
#define PDL_SSV2_0 1          /*   proposal sigma * v^2 */
#define PDL_Q0 2              /*   proposal q = 2 pi / d0 (line active while q0 < 2k) */
#define PDL_D 3               /*   model d */
#define PDL_D0 4              /*   proposal d */
#define PDL_DLS 5             /*   d log(sigma) / d d (model) */
#define PDL_Q2 6              /*   model q^2 */
/* SAbsorber (no reference counterpart) */
#define AB_SIGMA_A 0          /* 1/cm */
/* SDetectorTubes (no reference counterpart) */
#define DT_THMIN 0            /* degrees */
#define DT_THMAX 1
#define DT_YMIN 2             /* m */
#define DT_YMAX 3
#define DT_TMIN 4             /* microseconds */
#define DT_TMAX 5
#define DTI_NTUBE 0
#define DTI_NPIX 1
#define DTI_NTOF 2
#define DTI_RESTORE 3
/* SMCPLIn: particle records in the canonical layout MCR_* (floats, or
 * doubles when MCI_DOUBLE is set), converted from the file on the host
 * without arithmetic and stored field-major (field k of particle i at
 * k * MCI_N + i) */
#define MCI_N 0               /* particles in the file (at most 2^32 - 1) */
#define MCI_DOUBLE 1          /* 1: records hold doubles (double-precision file) */
#define MCT_REC 0
#define MCR_PX 0              /* 0..2 polarisation */
#define MCR_X 3               /* 3..5 position, cm */
#define MCR_U 6               /* 6..8 packed direction (u1, u2, +-ekin in MeV) */
#define MCR_T 9               /* time, ms */
#define MCR_W 10              /* weight */
#define MCR_N 11
/* SMCPLOut */
#define MCO_SLOT 0            /* index among the plan's SMCPLOut components */
#define MCO_STANDARD 1        /* 1: MCPL's own direction packing instead of the reference's */
/* SRescal */
#define RS_T 0                /* 0..2 target (sample frame) */
#define RS_E0 3               /* meV */
#define RS_DE 4               /* meV */
#define RS_FOCUS_R 5          /* m */
#define RSI_SLOT 0            /* index among the plan's SRescal components */
#define RSI_CAP_LO 6          /* set by the library: event buffer stride (histories), */
#define RSI_CAP_HI 7          /* low and high words; the buffer itself is DPart.aux */
#define RS_NEV 5              /* event fields: qx qy qz dE p */
/* SPolariser, SAnalyser */
#define PO_P 0                /* 0..2 polarisation vector */
/* SLinearCollimator (slopes = tan of the divergences, as collimator_lin.py forms them) */
#define CO_LEN 0              /* m (passed, unused by the kernel) */
#define CO_SLOPE_H 1
#define CO_SLOPE_V 2
#define CO_TRANS 3

#endif
