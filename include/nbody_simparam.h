/*
 * nbody_simparam.h - C entry point of libnbody_sim's SimParam parser, for
 * callers that are not C++ (ctypes, cgo) but want the reference's positional
 * command line (SimParam::parseArgs, src/sim_param.cpp:40-67).
 */
#ifndef NBODY_SIMPARAM_H
#define NBODY_SIMPARAM_H

#include "nbody_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Fills *out with SimParam() defaults overridden by argv[1..9], exactly as
 * SimParam::parseArgs reads them (argv[0] is the program name).  Returns
 * NB_OK, or NB_ERR_INVALID_ARG for a calcMethod other than BRANCH /
 * PREDICATED (where the reference throws std::invalid_argument). */
int nb_simparam_parse(int argc, char **argv, nb_params *out);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_SIMPARAM_H */
