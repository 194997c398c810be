// minim.h -- the library's umbrella header, as the reference's include/minim.h:8-23:
// State, Minimiser (Lbfgs, Fire, GradDescent), Potential and Communicator, and
// the potentials LjNd, BarAndHinge, PhaseField and PhaseFieldUnstructured.
// The reference's Anneal minimiser and its utils/mpi.h and utils/print.h are
// not part of this library.
#ifndef MINIM_AMD_MINIM_H
#define MINIM_AMD_MINIM_H

#include "Communicator.h"
#include "Minimiser.h"
#include "Potential.h"
#include "State.h"

#include "BarAndHinge.h"
#include "LjNd.h"
#include "PhaseField.h"
#include "PhaseFieldUnstructured.h"

#endif
