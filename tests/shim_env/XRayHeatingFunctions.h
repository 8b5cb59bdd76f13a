/* Stand-in for src/core/XRayHeatingFunctions.h: the per-snapshot stores ComputeTs keeps.  Test only. */
#ifndef MXR_SHIM_ENV_XRAY_H
#define MXR_SHIM_ENV_XRAY_H
extern double* stored_fcoll;
extern double* stored_XrayEmissivity_HMXB;
extern double* stored_XrayEmissivity_hard;
extern double* stored_XrayEmissivity_soft;
#endif
