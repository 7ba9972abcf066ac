"""ctypes binding of libllg_amd.so (the C ABI declared in include/llg/*.h).

Loading fails loudly when the library is missing: the HIP path is the
product, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libllg_amd.so"

# ---------------------------------------------------------------- constants
EXEC_YMSK, EXEC_COMPACT = 0, 1
ANISO_LINEAR, ANISO_CUBIC = 0, 1
NORM_NONE, NORM_EXACT, NORM_REGULARIZED = 0, 1, 2
BC_PERIODIC, BC_NEUMANN = 0, 1
DEMAG_OFF, DEMAG_PERIODIC, DEMAG_PADDED = 0, 1, 2
GATE_NONE, GATE_MASK, GATE_WINDOW = 0, 1, 2
PC_ALWAYS, PC_JOK = 0, 1
TERMS_RHS, TERMS_JV, TERMS_PC = 1, 2, 4
PREC_NONE, PREC_LEFT, PREC_RIGHT, PREC_BOTH = 0, 1, 2, 3
MODIFIED_GS, CLASSICAL_GS = 1, 2


class llg_terms(C.Structure):
    _fields_ = [
        ("che", C.c_double), ("chk", C.c_double), ("cha", C.c_double), ("chb", C.c_double),
        ("msk", C.c_double * 3), ("nsk", C.c_double * 3), ("happ", C.c_double * 3),
        ("aniso", C.c_int), ("cell_axes", C.c_int), ("window", C.c_int), ("gate", C.c_int),
    ]


class llg_config(C.Structure):
    _fields_ = [
        ("ng", C.c_int), ("ny", C.c_int), ("exec_", C.c_int), ("boundary", C.c_int),
        ("exch_self", C.c_int), ("normalize", C.c_int), ("demag", C.c_int),
        ("demag_fp32", C.c_int), ("pc_mode", C.c_int), ("pad0", C.c_int),
        ("demag_strength", C.c_double), ("demag_thick", C.c_double),
        ("alpha", C.c_double), ("chg", C.c_double),
        ("rhs", llg_terms), ("jv", llg_terms), ("pc", llg_terms),
    ]


class llgcv_stats(C.Structure):
    _fields_ = [
        ("nst", C.c_long), ("nfe", C.c_long), ("nsetups", C.c_long), ("nni", C.c_long),
        ("ncfn", C.c_long), ("netf", C.c_long), ("nli", C.c_long), ("nlcf", C.c_long),
        ("njvevals", C.c_long), ("npe", C.c_long), ("nps", C.c_long), ("nge", C.c_long),
        ("nfeLS", C.c_long), ("qlast", C.c_int), ("qcur", C.c_int),
        ("hlast", C.c_double), ("hcur", C.c_double), ("tcur", C.c_double), ("h0u", C.c_double),
        ("nsyncs", C.c_long),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class llg_run_opts(C.Structure):
    _fields_ = [
        ("t0", C.c_double), ("t_out", C.c_double), ("rtol", C.c_double), ("atol", C.c_double),
        ("maxl", C.c_int), ("gstype", C.c_int), ("max_order", C.c_int), ("fused", C.c_int),
        ("mxstep", C.c_long), ("use_jtv", C.c_int), ("use_prec", C.c_int),
        ("pretype", C.c_int), ("pad", C.c_int),
    ]


class llg_run_result(C.Structure):
    _fields_ = [
        ("st", llgcv_stats), ("wall_s", C.c_double), ("gpu_ms", C.c_double),
        ("sim_per_wall", C.c_double), ("t_reached", C.c_double), ("nout", C.c_long),
        ("retval", C.c_int), ("pad", C.c_int),
    ]


class llg_grain(C.Structure):
    _fields_ = [
        ("sx", C.c_double), ("sy", C.c_double), ("ax", C.c_double), ("ay", C.c_double),
        ("az", C.c_double), ("nx", C.c_double), ("ny", C.c_double), ("nz", C.c_double),
        ("dead", C.c_int), ("pad", C.c_int),
    ]


class llg_instance(C.Structure):
    _fields_ = [
        ("cfg", llg_config), ("active", C.POINTER(C.c_ubyte)), ("w", C.POINTER(C.c_double)),
        ("cell_msk", C.POINTER(C.c_double)), ("cell_nsk", C.POINTER(C.c_double)),
        ("y0", C.POINTER(C.c_double)),
    ]


class llg_res_in(C.Structure):
    _fields_ = [("zn1", C.c_void_p), ("ycor", C.c_void_p),
                ("c0", C.c_double), ("c1", C.c_double), ("c2", C.c_double)]


class llg_ensemble_result(C.Structure):
    _fields_ = [
        ("wall_s", C.c_double), ("inst_time", C.c_double), ("nst", C.c_long),
        ("nfe", C.c_long), ("nli", C.c_long), ("nni", C.c_long), ("failed", C.c_int),
        ("pad", C.c_int),
    ]


DP = C.POINTER(C.c_double)
UCP = C.POINTER(C.c_ubyte)
IP = C.POINTER(C.c_int)
VP = C.c_void_p

# name: (restype, argtypes) for every exported entry point of include/llg/*.h
SIGNATURES = {
    # nvector_hip.h
    "N_VNew_Hip": (VP, [C.c_int64, VP]),
    "N_VNewWithStream_Hip": (VP, [C.c_int64, VP, VP]),
    "N_VNewEmpty_Hip": (VP, [VP]),
    "N_VMake_Hip": (VP, [C.c_int64, DP, VP, VP]),
    "N_VGetDeviceArrayPointer_Hip": (DP, [VP]),
    "N_VGetHostArrayPointer_Hip": (DP, [VP]),
    "N_VCopyToDevice_Hip": (None, [VP]),
    "N_VCopyFromDevice_Hip": (None, [VP]),
    "N_VGetStream_Hip": (VP, [VP]),
    "N_VGetLength_Hip": (C.c_int64, [VP]),
    "N_VEnableFusedOps_Hip": (C.c_int, [VP, C.c_int]),
    "N_VSetFromHost_Hip": (C.c_int, [VP, DP]),
    "N_VGetToHost_Hip": (C.c_int, [VP, DP]),
    "NVHip_DefaultStream": (VP, []),
    "NVHip_NumSyncs": (C.c_long, []),
    "NVHip_SetReduceConfig": (C.c_int, [C.c_int, C.c_int]),
    "NVHip_SetSyncMode": (C.c_int, [C.c_int]),
    "N_VGetVectorID_Hip": (C.c_int, [VP]),
    "N_VClone_Hip": (VP, [VP]),
    "N_VCloneEmpty_Hip": (VP, [VP]),
    "N_VDestroy_Hip": (None, [VP]),
    "N_VSpace_Hip": (None, [VP, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "N_VLinearSum_Hip": (None, [C.c_double, VP, C.c_double, VP, VP]),
    "N_VConst_Hip": (None, [C.c_double, VP]),
    "N_VProd_Hip": (None, [VP, VP, VP]),
    "N_VDiv_Hip": (None, [VP, VP, VP]),
    "N_VScale_Hip": (None, [C.c_double, VP, VP]),
    "N_VAbs_Hip": (None, [VP, VP]),
    "N_VInv_Hip": (None, [VP, VP]),
    "N_VAddConst_Hip": (None, [VP, C.c_double, VP]),
    "N_VDotProd_Hip": (C.c_double, [VP, VP]),
    "N_VMaxNorm_Hip": (C.c_double, [VP]),
    "N_VWrmsNorm_Hip": (C.c_double, [VP, VP]),
    "N_VWrmsNormMask_Hip": (C.c_double, [VP, VP, VP]),
    "N_VMin_Hip": (C.c_double, [VP]),
    "N_VWL2Norm_Hip": (C.c_double, [VP, VP]),
    "N_VL1Norm_Hip": (C.c_double, [VP]),
    "N_VCompare_Hip": (None, [C.c_double, VP, VP]),
    "N_VInvTest_Hip": (C.c_int, [VP, VP]),
    "N_VConstrMask_Hip": (C.c_int, [VP, VP, VP]),
    "N_VMinQuotient_Hip": (C.c_double, [VP, VP]),
    "N_VWSqrSumLocal_Hip": (C.c_double, [VP, VP]),
    "N_VWSqrSumMaskLocal_Hip": (C.c_double, [VP, VP, VP]),
    "N_VLinearCombination_Hip": (C.c_int, [C.c_int, DP, C.POINTER(VP), VP]),
    "N_VScaleAddMulti_Hip": (C.c_int, [C.c_int, DP, VP, C.POINTER(VP), C.POINTER(VP)]),
    "N_VDotProdMulti_Hip": (C.c_int, [C.c_int, VP, C.POINTER(VP), DP]),
    "N_VDotProdMultiNoAlloc_Hip": (C.c_int, [C.c_int, VP, C.POINTER(VP), DP]),
    "N_VLinearSumVectorArray_Hip": (C.c_int, [C.c_int, C.c_double, C.POINTER(VP), C.c_double,
                                              C.POINTER(VP), C.POINTER(VP)]),
    "N_VScaleVectorArray_Hip": (C.c_int, [C.c_int, DP, C.POINTER(VP), C.POINTER(VP)]),
    "N_VConstVectorArray_Hip": (C.c_int, [C.c_int, C.c_double, C.POINTER(VP)]),
    "N_VWrmsNormVectorArray_Hip": (C.c_int, [C.c_int, C.POINTER(VP), C.POINTER(VP), DP]),
    "N_VWrmsNormMaskVectorArray_Hip": (C.c_int, [C.c_int, C.POINTER(VP), C.POINTER(VP), VP, DP]),
    "NVHip_Pascal": (C.c_int, [C.c_int, C.POINTER(VP), C.c_double]),
    "NVHip_EwtSet": (C.c_int, [VP, C.c_double, C.c_double, VP, VP]),
    "NVHip_PredictInit": (C.c_int, [C.c_int, C.POINTER(VP), VP, VP, C.c_double, C.c_double,
                                    VP, VP]),
    "NVHip_WrmsNormPair": (C.c_int, [VP, C.c_double, VP, C.c_double, VP, VP, DP]),
    "NVHip_PredictComplete": (C.c_int, [C.c_int, C.POINTER(VP), VP, DP, DP, VP, C.c_double,
                                        C.c_double, VP, VP]),
    "NVHip_PredictCompleteSave": (C.c_int, [C.c_int, C.POINTER(VP), VP, DP, DP, VP, C.c_double,
                                        C.c_double, VP, VP, VP]),
    "NVHip_WrmsNormPairComp": (C.c_int, [VP, C.c_double, VP, C.c_double, VP, C.c_double, VP,
                                         VP, DP]),
    "NVHip_WrmsNormPairCompQueue": (C.c_int, [VP, C.c_double, VP, C.c_double, VP, C.c_double,
                                              VP, VP, C.POINTER(C.c_uint)]),
    "NVHip_WrmsNormPairWait": (C.c_int, [VP, C.c_uint, DP]),
    "NVHip_ScaleDiv": (C.c_int, [C.c_double, VP, VP, VP, VP]),
    "NVHip_ProdDotMulti": (C.c_int, [VP, VP, VP, C.c_int, C.POINTER(VP), DP]),
    "NVHip_LinCombDot": (C.c_int, [C.c_int, DP, C.POINTER(VP), VP, DP]),
    "NVHip_GSProject": (C.c_int, [VP, VP, VP, C.c_int, C.POINTER(VP), DP, DP]),
    "NVHip_GSProjectScaled": (C.c_int, [VP, C.c_int, C.POINTER(VP), C.c_int, C.c_double, DP, DP]),
    "NVHip_SpgmrSolve": (C.c_int, [VP, VP, VP, C.POINTER(VP), C.c_int, VP, VP, C.c_double, IP, IP,
                                   DP, DP]),
    "NVHip_LinCombDiv": (C.c_int, [C.c_int, DP, C.POINTER(VP), VP, VP]),
    "NVHip_ReducePartials": (DP, [VP, IP]),
    "NVHip_ReduceFinish": (C.c_int, [VP, C.c_int, C.c_int, DP]),
    "NVHip_ReduceFinishAsync": (C.c_int, [VP, C.c_int, C.c_int, VP, C.POINTER(C.c_uint)]),
    "NVHip_ReduceWait": (C.c_int, [VP, C.c_uint, C.c_int, DP]),
    "NVHip_GateOpen": (C.c_int, [VP, C.POINTER(VP), C.POINTER(VP)]),
    "NVHip_GateRelease": (C.c_int, [VP, C.c_int, DP, C.c_int]),
    "NVHip_LagBegin": (C.c_int, [VP, C.c_int, C.c_double, C.c_double, C.c_double,
                                 C.POINTER(VP), C.POINTER(VP), C.POINTER(VP)]),
    "NVHip_LagStep": (C.c_int, [VP, C.c_int, C.c_int, C.c_int, C.c_int]),
    "NVHip_LagNext": (C.c_int, [VP]),
    "NVHip_LagUpdate": (C.c_int, [VP, C.c_int, C.POINTER(VP), VP, VP, VP, VP, VP, VP,
                                  C.POINTER(C.c_uint)]),
    "NVHip_LagUpdateWait": (C.c_int, [VP, C.c_uint, DP]),
    "NVHip_LagProjNorm": (C.c_int, [VP, C.c_int, C.POINTER(VP)]),
    "NVHip_LagWait": (C.c_int, [VP, VP]),
    "NVHip_ProjNorm": (C.c_int, [C.c_int, C.c_double, DP, VP, C.POINTER(VP), VP, DP]),
    "NVHip_NewtonUpdate": (C.c_int, [C.c_int, VP, VP, VP, VP, VP, DP, DP]),
    "NVHip_NewtonUpdateComb": (C.c_int, [VP, C.c_int, DP, C.POINTER(VP), VP, VP, VP, VP, DP]),
    "FusedNVec_Init": (None, [VP]),
    "FusedNVec_FreePool": (None, []),
    # demag_fft.h
    "Demag_Init": (VP, [C.c_int, C.c_int, C.c_double, C.c_double]),
    "Demag_InitPadded": (VP, [C.c_int, C.c_int, C.c_double, C.c_double]),
    "Demag_Apply": (None, [VP, DP, DP]),
    "Demag_ApplyWindowed": (None, [VP, DP, DP, DP]),
    "Demag_SetStream": (None, [VP, VP]),
    "Demag_ApplyBatch": (C.c_int, [C.c_int, C.POINTER(VP), C.POINTER(VP), C.POINTER(VP),
                                   C.POINTER(VP), VP]),
    "Demag_GetSelfCoupling": (None, [VP, DP, DP, DP]),
    "Demag_Destroy": (None, [VP]),
    "Demag_SetLdsPadding": (C.c_int, [VP, C.c_int, C.c_int]),
    "Demag_SetSpecialized": (C.c_int, [VP, C.c_int]),
    "Demag_SetWorkgroups": (C.c_int, [VP, C.c_int, C.c_int]),
    "Demag_Autotune": (C.c_int, [VP, C.c_int, IP]),
    "Demag_SetFftKernels": (C.c_int, [VP, C.c_int, C.c_int]),
    "Demag_GetFftKernels": (C.c_int, [VP, IP, IP]),
    "Demag_SetPrecision": (C.c_int, [VP, C.c_int]),
    "Demag_GetPlanShape": (C.c_int, [VP, IP, IP, IP]),
    "Demag_GetSpectrum": (C.c_int, [VP, DP]),
    "Demag_LastError": (C.c_char_p, []),
    "Demag_NewellTensor": (C.c_int, [C.c_double, C.c_int, C.c_int, DP, C.c_int]),
    # llg_plugin.h
    "LLG_Create": (VP, [C.POINTER(llg_config), UCP, DP, DP, DP, VP]),
    "LLG_Destroy": (None, [VP]),
    "LLG_LastError": (C.c_char_p, []),
    "LLG_GetCounts": (C.c_int, [VP, IP, IP, IP]),
    "LLG_GetMask": (C.c_int, [VP, DP]),
    "LLG_GetActiveIds": (C.c_int, [VP, IP, IP]),
    "LLG_GetHdemag": (C.c_int, [VP, DP]),
    "LLG_GetPinv": (C.c_int, [VP, DP]),
    "LLG_GetDemag": (VP, [VP]),
    "LLG_GetStream": (VP, [VP]),
    "LLG_GetCallCounts": (C.c_int, [VP, C.POINTER(C.c_long), C.POINTER(C.c_long),
                                    C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    "LLG_ProfileStart": (C.c_int, [C.c_int, C.c_int]),
    "LLG_ProfileSetEvery": (C.c_int, [C.c_int]),
    "LLG_ProfileStop": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    "LLG_EventCreate": (VP, []),
    "LLG_EventRecord": (C.c_int, [VP, VP]),
    "LLG_EventElapsedMs": (C.c_double, [VP, VP]),
    "LLG_EventDestroy": (None, [VP]),
    "LLG_DeviceSynchronize": (C.c_int, []),
    "LLG_Rhs": (C.c_int, [C.c_double, VP, VP, VP]),
    "LLG_RhsResidual": (C.c_int, [C.c_double, VP, DP, VP, VP, VP, VP]),
    "JtvProduct": (C.c_int, [VP, VP, C.c_double, VP, VP, VP, VP]),
    "LLG_Atimes": (C.c_int, [VP, VP, C.c_double, C.c_double, VP, VP, VP]),
    "LLG_Krylov": (C.c_int, [C.c_double, VP, VP, VP, C.c_double, C.c_double, VP, VP, VP]),
    "LLG_SpgmrSolve": (C.c_int, [VP, C.POINTER(VP), C.c_int, VP, VP, C.c_double, C.c_double,
                                 C.c_double, VP, VP, VP, IP, IP, DP, DP]),
    "LLG_Arnoldi": (C.c_int, [C.c_int, C.c_int, C.c_double, DP, VP, C.POINTER(VP), VP, VP,
                              C.c_double, C.c_double, VP, VP, VP, DP]),
    "LLG_ArnoldiQueue": (C.c_int, [C.c_int, C.c_int, C.c_double, DP, VP, VP, VP, C.POINTER(VP),
                                   VP, VP, C.c_double, C.c_double, VP, VP, VP,
                                   C.POINTER(C.c_uint)]),
    "LLG_ArnoldiLag": (C.c_int, [C.c_int, C.c_int, C.c_int, VP, VP, VP, C.POINTER(VP), VP,
                                 VP, C.c_double, C.c_double, VP, VP, VP]),
    "LLG_ArnoldiFirst": (C.c_int, [C.c_int, DP, C.POINTER(VP), C.POINTER(VP), VP, VP,
                                   C.c_double, C.c_double, VP, VP, VP, DP]),
    "PrecondSetup": (C.c_int, [C.c_double, VP, VP, C.c_int, C.POINTER(C.c_int), C.c_double, VP]),
    "PrecondSolve": (C.c_int, [C.c_double, VP, VP, VP, VP, C.c_double, C.c_double, C.c_int, VP]),
    "Precond_Create": (VP, [C.c_int, C.c_int, C.c_int]),
    "Precond_Destroy": (None, [VP]),
    # llg_geometry.h
    "LLG_GeomSquareHole": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, UCP]),
    "LLG_GeomRectangle": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, UCP]),
    "LLG_GeomRing": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, UCP]),
    "LLG_GeomEllipse": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, UCP]),
    "LLG_GeomPolycrystal": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint, C.c_double,
                                      C.c_double, C.POINTER(llg_grain), IP, UCP, DP, DP, DP]),
    "LLG_ICUniform": (C.c_int, [C.c_int, C.c_int, UCP, C.c_double, C.c_double, C.c_double, DP]),
    "LLG_ICTwoDomain": (C.c_int, [C.c_int, C.c_int, UCP, C.c_int, C.c_double, DP]),
    "LLG_ICStripes": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint, DP]),
    "LLG_ICGrainBumps": (C.c_int, [C.c_int, C.c_int, C.POINTER(llg_grain), IP, DP, C.c_double,
                                   C.c_uint, DP]),
    "LLG_ICTiltedZ": (C.c_int, [C.c_int, C.c_int, UCP, C.c_double, C.c_uint64, DP]),
    "LLG_ICSplitX": (C.c_int, [C.c_int, C.c_int, UCP, C.c_double, C.c_uint64, DP]),
    "LLG_ICRandomSphere": (C.c_int, [C.c_int, C.c_int, UCP, C.c_uint64, DP]),
    # llg_output.h
    "LLG_FrameWriterOpen": (VP, [C.c_char_p]),
    "LLG_FrameWriterWrite": (C.c_int, [VP, C.c_double, C.c_int, C.c_int, C.c_int, DP, DP, C.c_int]),
    "LLG_FrameWriterClose": (C.c_int, [VP]),
    "LLG_ShouldWriteFrame": (C.c_int, [C.c_long, C.c_double, C.c_double, C.c_int, C.c_int]),
    "LLG_WriteMaskFile": (C.c_int, [C.c_char_p, C.c_int, C.c_int, DP]),
    "LLG_WriteGrainIdFile": (C.c_int, [C.c_char_p, C.c_int, C.c_int, IP]),
    # llg_cvode.h
    "LLGCVodeCreate": (VP, [C.c_int, VP]),
    "LLGCVodeSetFusedOps": (C.c_int, [VP, VP]),
    # learned_pc.h
    "LLG_LpcCreate": (VP, [VP, C.c_int]),
    "LLG_LpcDestroy": (None, [VP]),
    "LLG_LpcLastError": (C.c_char_p, []),
    "LLG_LpcNumWeights": (C.c_int, [VP]),
    "LLG_LpcSetWeights": (C.c_int, [VP, C.POINTER(C.c_float)]),
    "LLG_LpcGetWeights": (C.c_int, [VP, C.POINTER(C.c_float)]),
    "LLG_LpcGetGrad": (C.c_int, [VP, C.POINTER(C.c_float)]),
    "LLG_LpcForward": (C.c_int, [VP, DP, DP, C.c_double, C.c_double, C.POINTER(C.c_float)]),
    "LLG_LpcLossGrad": (C.c_int, [VP, DP, DP, DP, C.c_double, C.c_double, DP]),
    "LLG_LpcAdamStep": (C.c_int, [VP, C.c_float]),
    "LLG_LpcTrainStep": (C.c_int, [VP, DP, DP, DP, C.c_double, C.c_double, C.c_float, DP]),
    "LLG_LpcCorrect": (C.c_int, [VP, DP, C.c_double, DP]),
    "LLG_LpcApply": (C.c_int, [VP, DP, C.c_double, DP, DP]),
    "LLG_LpcInputBuffer": (DP, [VP]),
    "LLG_LpcSetFused": (C.c_int, [VP, C.c_int]),
    "LLG_SetLearnedPrecond": (C.c_int, [VP, VP]),
    "LLG_GetGrid": (C.c_int, [VP, IP, IP]),
    "LLG_SetAppliedField": (C.c_int, [VP, C.c_int, DP]),
    "LLG_GetAppliedField": (C.c_int, [VP, C.c_int, DP]),
    "LLG_MeanMagnetization": (C.c_int, [VP, VP, DP]),
    "LLG_SetStencilBlock": (C.c_int, [VP, C.c_int, C.c_int]),
    "LLG_Autotune": (C.c_int, [VP, C.c_int, IP]),
    "LLG_GetActiveDevice": (VP, [VP]),
    # llg_ensemble.h
    "LLG_EnsembleCreate": (VP, [C.c_int, C.POINTER(llg_instance), C.POINTER(llg_run_opts),
                                C.c_int]),
    "LLG_EnsembleAdvance": (C.c_int, [VP, C.c_double, C.POINTER(llg_ensemble_result)]),
    "LLG_EnsembleSize": (C.c_int, [VP]),
    "LLG_EnsembleGetState": (C.c_int, [VP, C.c_int, DP]),
    "LLG_EnsembleGetResult": (C.c_int, [VP, C.c_int, C.POINTER(llg_run_result)]),
    "LLG_EnsembleDestroy": (None, [VP]),
    "LLG_EnsembleSetBatched": (C.c_int, [VP, C.c_int]),
    "LLG_EnsembleBatchStats": (C.c_int, [VP, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    # llg_batch.h
    "LLG_BatchCreate": (VP, [C.c_int, C.POINTER(VP)]),
    "LLG_BatchSize": (C.c_int, [VP]),
    "LLG_RhsBatch": (C.c_int, [VP, C.POINTER(VP), C.POINTER(VP), C.POINTER(VP), VP]),
    "LLG_JtvBatch": (C.c_int, [VP, C.POINTER(VP), C.POINTER(VP), C.POINTER(VP), VP]),
    "LLG_BatchDestroy": (None, [VP]),
    "LLGCVodeInit": (C.c_int, [VP, VP, C.c_double, VP]),
    "LLGCVodeReInit": (C.c_int, [VP, C.c_double, VP]),
    "LLGCVodeSetUserData": (C.c_int, [VP, VP]),
    "LLGCVodeSStolerances": (C.c_int, [VP, C.c_double, C.c_double]),
    "LLGCVodeSVtolerances": (C.c_int, [VP, C.c_double, VP]),
    "LLGCVodeSetMaxOrd": (C.c_int, [VP, C.c_int]),
    "LLGCVodeSetMaxNumSteps": (C.c_int, [VP, C.c_long]),
    "LLGCVodeSetInitStep": (C.c_int, [VP, C.c_double]),
    "LLGCVodeSetMaxStep": (C.c_int, [VP, C.c_double]),
    "LLGCVodeSetMinStep": (C.c_int, [VP, C.c_double]),
    "LLGCVodeSetStopTime": (C.c_int, [VP, C.c_double]),
    "LLGCVodeSetLinearSolverSPGMR": (C.c_int, [VP, C.c_int, C.c_int, C.c_int]),
    "LLGCVodeSetEpsLin": (C.c_int, [VP, C.c_double]),
    "LLGCVodeSetJacTimes": (C.c_int, [VP, VP, VP]),
    "LLGCVodeSetPreconditioner": (C.c_int, [VP, VP, VP]),
    "LLGCVode": (C.c_int, [VP, C.c_double, VP, C.POINTER(C.c_double), C.c_int]),
    "LLGCVodeGetDky": (C.c_int, [VP, C.c_double, C.c_int, VP]),
    "LLGCVodeGetStats": (C.c_int, [VP, C.POINTER(llgcv_stats)]),
    "LLGCVodeFree": (None, [C.POINTER(VP)]),
    # llg_driver.h
    "LLG_DefaultRunOpts": (None, [C.POINTER(llg_run_opts)]),
    "LLG_RunCreate": (VP, [VP, C.POINTER(llg_run_opts), DP]),
    "LLG_RunAdvance": (C.c_int, [VP, C.c_double, C.POINTER(llg_run_result)]),
    "LLG_RunGetState": (C.c_int, [VP, DP]),
    "LLG_RunStage": (C.c_int, [VP, C.c_int, DP, C.c_double, C.POINTER(llg_run_result), DP]),
    "LLG_RunRamp": (C.c_int, [VP, C.c_int, C.c_int, DP, C.c_double, DP,
                              C.POINTER(llg_run_result)]),
    "LLG_RunDestroy": (None, [VP]),
}

_lib = None


def load(path: str | Path | None = None):
    """Load libllg_amd.so and attach prototypes (raises if it is missing)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"native library {p} is missing: build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def fn_addr(lib, name):
    """Address of an exported C function (for callback registration)."""
    return C.cast(getattr(lib, name), C.c_void_p).value
