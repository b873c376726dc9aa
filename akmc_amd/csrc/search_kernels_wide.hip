// The walker state machines of search_kernels.hip in 1024-thread
// workgroups (akmc_launch_search_logic_wide), for batches whose walkers are
// large enough that the strided loops of a 256-thread workgroup dominate a
// micro-step: the full relax of a cell of tens of thousands of atoms.
#define AKMC_LOGIC_BLOCK 1024
#include "search_kernels.hip"
