/*
 * radchan_param_shim.h -- the reference's compile-time configuration,
 * carried over to the library at load time.
 *
 * RadChan configures its solver with the macros of src/param/param.h and
 * src/param/NarrowBand.h.  A DNS build that switches to this library keeps
 * those two headers and adds one small C++ file, in this order:
 *
 *     #include "radchan_param_shim.h"      // this file, first
 *     #include "param.h"                   // the reference's, unchanged
 *     #include "NarrowBand.h"              // idem
 *     #include "radchan_param_shim_def.h"  // the load-time constructor
 *
 * The constructor passes every macro value to rc_params_compiled and the
 * result to rc_set_params, so mc_gpu_/get_results_ run with exactly the
 * compiled-in configuration.  The macro names equal the rc_params field
 * names, which is why radchan.h has to be seen before param.h and why the
 * constructor never names a field.
 *
 * One GPU owns one node's slab: p_col keeps its meaning (param.h:9); the
 * reference's num_gpu/num_streams/blockdim/nblocks (param.h:8,52-54) only
 * shaped the CUDA launches and have no counterpart here.
 */
#ifndef RADCHAN_PARAM_SHIM_H
#define RADCHAN_PARAM_SHIM_H
#include <stdio.h>
#include "radchan.h"
#endif
