// Drop-in for the reference header of the same name: see rsa_mpi_amd.hpp.
#pragma once
#include "rsa_mpi_amd.hpp"
// as the reference header does (global scope)
using namespace sac_de_billes;
