// globalVars.h -- physical constants of the wave-CAIPI trajectory
// (values of the reference's globalVars.h:3-4: the truncated pi and the
// gyromagnetic ratio over 2 pi of 1H in Hz/T).
#pragma once

const double PI = 3.14159265359;
const double gammaBar = 4.257748059639612e7;
