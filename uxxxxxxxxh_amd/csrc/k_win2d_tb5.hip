// 5 timesteps per launch for the 2D stars (win2d_tbk with K = 5).
#include "k_win2d_tbk.h"

LaunchFn uhs_find_win2d_tb5(int prog, int dtype, int vec) { return win2d_tbk_find<5>(prog, dtype, vec); }
