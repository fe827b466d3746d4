// Drop-in for the reference's operator_paraview.hxx: rsa_paraview::paraview
// (VTK XML PolyData output of the placed spheres).
#pragma once
#include "rsa_mpi_amd.hpp"
