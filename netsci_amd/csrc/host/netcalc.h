// netcalc:: — the reference's compute API on top of libnsx.
//
// Signatures follow netcalc/include/{mutual_information,generalized_correlation,
// hedetniemi,platform}.h. The pair loops return `platform` (0 GPU, 1 CPU) as
// the reference does (generalized_correlation.cpp:101,172) and throw
// std::runtime_error on an invalid platform (:83,167); hedetniemi throws
// std::invalid_argument (hedetniemi.cpp:35). Failures of the GPU engine
// (no device, bad arguments) throw std::runtime_error with libnsx's message;
// nothing falls back to the CPU.
#pragma once

#include <string>

#include "cuarray.h"

namespace netcalc {

const int GPU_PLATFORM = 0;  // platform.h:9
const int CPU_PLATFORM = 1;  // platform.h:10

// mutual_information.h:34-43 / mutual_information.cpp:103-172.
int mutualInformation(CuArray<float> *X, CuArray<float> *I, CuArray<int> *ab, int k, int n,
                      int xd, int d, int platform);
// mutual_information.h:66-77 / mutual_information.cpp:8-101.
int mutualInformation(CuArray<float> *X, CuArray<float> *I, CuArray<int> *ab, int k, int n,
                      int xd, int d, int platform, int checkpointFrequency,
                      std::string checkpointFileName);
// mutual_information.h:91-120.
float mutualInformationGpu(CuArray<float> *Xa, CuArray<float> *Xb, int k, int n, int xd, int d);
float mutualInformationCpu(CuArray<float> *Xa, CuArray<float> *Xb, int k, int n, int xd, int d);

// generalized_correlation.h:26-106 / generalized_correlation.cpp:8-173.
int generalizedCorrelation(CuArray<float> *X, CuArray<float> *R, CuArray<int> *ab, int k, int n,
                           int xd, int d, int platform);
int generalizedCorrelation(CuArray<float> *X, CuArray<float> *R, CuArray<int> *ab, int k, int n,
                           int xd, int d, int platform, int checkpointFrequency,
                           std::string checkpointFileName);
float generalizedCorrelationGpu(CuArray<float> *Xa, CuArray<float> *Xb, int k, int n, int xd,
                                int d);
float generalizedCorrelationCpu(CuArray<float> *Xa, CuArray<float> *Xb, int k, int n, int xd,
                                int d);

// mutual_information.h:132-136 / mutual_information.cpp:174-201.
void generateRestartAbFromCheckpointFile(CuArray<int> *ab, CuArray<int> *restartAb,
                                         const std::string &checkpointFileName);

// hedetniemi.h:31-180 / hedetniemi.cpp:10-108 / hedetniemi_gpu.cu:109-261.
void hedetniemiAllShortestPaths(CuArray<float> *A, CuArray<float> *H, CuArray<int> *paths,
                                int maxPathLength, float tolerance, int platform);
void hedetniemiAllShortestPathsGpu(CuArray<float> *A, CuArray<float> *H, CuArray<int> *paths,
                                   float tolerance, int maxPathLength);
void hedetniemiAllShortestPathsCpu(CuArray<float> *A, CuArray<float> *H, CuArray<int> *paths,
                                   float tolerance, int maxPathLength);
void hedetniemiAllShortestPathLengths(CuArray<float> *A, CuArray<float> *H, int maxPathLength,
                                      int platform);
void hedetniemiAllShortestPathLengthsGpu(CuArray<float> *A, CuArray<float> *H,
                                         int maxPathLength);
void correlationToAdjacency(CuArray<float> *A, CuArray<float> *C, int n, int platform);
void correlationToAdjacencyGpu(CuArray<float> *A, CuArray<float> *C, int n);
// math_utils.h:24-101: per-row mean and population standard deviation of an
// m x n matrix (m is recomputed as a.size() / n, as math_utils.cpp:14,32).
void mean(CuArray<float> *a, CuArray<float> *u, int m, int n, int platform);
void meanGpu(CuArray<float> *a, CuArray<float> *u, int m, int n);
void standardDeviation(CuArray<float> *a, CuArray<float> *u, CuArray<float> *sigma, int m, int n,
                       int platform);
void standardDeviationGpu(CuArray<float> *a, CuArray<float> *u, CuArray<float> *sigma, int m,
                          int n);
// hedetniemi.h:173-180 / hedetniemi.cpp:71-108, the reference's signature:
// the path i -> j as node indices in a new int[] of (*NUMPY_ARRAY_DIM1)[0]
// entries; *NUMPY_ARRAY_DIM1 is a new int[1]. Both are new[]-allocated, as
// the reference allocates them, and the caller releases them with delete[].
void recoverSingleShortestPath(int **NUMPY_ARRAY, int **NUMPY_ARRAY_DIM1, CuArray<int> *paths,
                               int maxPathLength, int i, int j);
// The same path as a vector (what the Python binding returns as an int32 array).
std::vector<int> recoverSingleShortestPath(CuArray<int> *paths, int maxPathLength, int i, int j);

// Bootstrap replicates as frame windows (no reference counterpart; each
// replicate equals the pair loop on that window's frames): R is re-initialised
// to nWindows x ab->m() GC (gc != 0) or MI values.
int windowedCorrelation(CuArray<float> *X, CuArray<float> *R, CuArray<int> *ab,
                        CuArray<int> *windows, int k, int nTotal, int d, int gc);

}  // namespace netcalc
