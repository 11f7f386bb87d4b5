// mtmd_output.h - trajectory and histogram file writers of mtmd_run (host only).
//
// The file formats are the reference's (Cmd/MolTwisterCmdMolDyn/MDLoop/MDLoop.cpp
// and Utilities/DCDFile.cpp), so that the files a `moldyn run` leaves behind read
// the same in the tools users already have (VMD for DCD, any XYZ reader).
#pragma once
#include <stdint.h>

// CMDLoop::appendToXYZFile (MDLoop.cpp:229-246): "<n>\r\nFrame <t>\r\n", then one
// line per atom: type name (or the type index when no names are given), x, y, z.
int mt_xyz_append(const char* path, int n, int t, const double* xyz, const int32_t* type,
                  const char* const* type_names, int n_types);

// CDCDFile::createDCDFileIfNotExists (DCDFile.cpp:670-697): the CHARMM main
// header when no file exists at path.
int mt_dcd_create_if_missing(const char* path, int n_steps, int stride, double timestep_fs, int n_atoms);

// CDCDFile::appendToDCDFile (DCDFile.cpp:699-729): one record, box widths as
// int (the reference narrows them), angles 0, coordinates as float.
int mt_dcd_append(const char* path, int n_atoms, const int box[3], const double* xyz);

// CMDLoop::storeMomentumDistribution (MDLoop.cpp:325-349): axis < 3 momentum
// bins (unit conversion Conv_P), axis 3 the thermostat momentum (Conv_Peta).
int mt_store_pdistr(const char* path, const int* hist, int nbins, double maxP, int axis);

// CMDLoop::storeVolumeDistribution (MDLoop.cpp:351-366).
int mt_store_vdistr(const char* path, const int* hist, int nbins, double maxV);
