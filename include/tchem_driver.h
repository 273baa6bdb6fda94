/* TChem_Driver coupling interface (libtchem_driver.so): the reference's
 * extern "C" surface for host models such as PartMC
 * (src/core/TChem_Driver.hpp:106-132), backed by the GPU aerosol integrator
 * of libtchem_atm.so (tchem_aerosolchem_run_device_batch).
 *
 * One process-wide model, as the reference: initialize() reads the gas
 * mechanism, the aerosol file and the solver settings, and allocates nBatch
 * host state vectors [density, pressure, temperature, gas species (active
 * then constant), particle species]; TChem_doTimestep(del_t) integrates every
 * batch entry from 0 to del_t on the current HIP device (TR-BDF2, FP64).
 *
 * Differences from the reference, all supersets: every batch entry is
 * integrated and copied back (the reference's doTimestep copies back entry 0
 * only, TChem_Driver.cpp:617-627); errors print a message to stderr and make
 * the call a no-op (getters return -1) instead of aborting.
 * TChem_getLastError() returns the last message.
 */
#ifndef TCHEM_DRIVER_H
#define TCHEM_DRIVER_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* TChem_Driver.hpp:106-109, TChem_Driver.cpp:86-127 */
void initialize(const char* gasFile, const char* aeroFile, const char* numericsFile, const int nBatch);
/* TChem_Driver.hpp:109, .cpp:132-137 */
void finalize(void);

/* number of gas species, active and constant (TChem_Driver.cpp:340-346) */
int TChem_getNumberOfSpecies(void);
/* state vector length: 3 + gas species + aerosol species x particles (.cpp:414-421) */
int TChem_getLengthOfStateVector(void);
/* TChem_getLengthOfStateVector (declared by the reference, TChem_Driver.hpp:123) */
int TChem_getStateVectorSize(void);
/* copy batch entry iBatch out of / into the driver's state (.cpp:281-311) */
void TChem_getStateVector(double* array, const int iBatch);
void TChem_setStateVector(double* array, const int iBatch);
/* all batch entries, row-major [nBatch][length] (TChem_Driver.hpp:111,124) */
void TChem_getAllStateVectorHost(double* view);
/* gas species name at *index; returns its length (.cpp:316-333) */
int TChem_getSpeciesName(int* index, char* result, const size_t buffer_size);

/* integrate every batch entry over del_t seconds (.cpp:448-632); the
 * reference takes a C++ reference, which has the ABI of a pointer */
#ifdef __cplusplus
void TChem_doTimestep(const double& del_t);
#else
void TChem_doTimestep(const double* del_t);
#endif

/* particles per batch entry (.cpp:426-435) */
int TChem_getNumberConcentrationVectorSize(void);
void TChem_setNumberConcentrationVector(double* array, const int iBatch);
/* aerosol species (.cpp:351-411) */
int TChem_getNumberOfAeroSpecies(void);
int TChem_getAerosolSpeciesName(int* index, char* result, const size_t buffer_size);
double TChem_getAerosolSpeciesDensity(int* index);
double TChem_getAerosolSpeciesMW(int* index);
double TChem_getAerosolSpeciesKappa(int* index);

/* no counterpart in the reference: the last error message ("" if none) */
const char* TChem_getLastError(void);

#ifdef __cplusplus
}
#endif

#endif
