// Constants.hh -- physical and numerical constants of the SWE step
// (same names and values as the reference's src/Constants.hh:4-6).
#ifndef SWE_HIP_CONSTANTS_HH
#define SWE_HIP_CONSTANTS_HH

const float g = 9.81f;
const float defaultDryTol = 0.1f;
const float defaultCflNumber = 0.4f;

#endif
