// internal.h -- shared declarations of the host side of libtfhe_mi355x.
#pragma once
#include <stdint.h>

#include <mutex>
#include <string>

#include <hip/hip_runtime.h>

#include "tfhe_mi355x.h"

namespace tfhe {

// last-error text for the flat ABI (thread-local)
void set_error(const std::string& msg);
// Aborts with a message; used by the void upstream entry points, which have no error channel
// (the reference host exits on any device error, kernel/src/include/host.h:13-18).
[[noreturn]] void fatal(const char* where, int code);

// flat key export of a secret/cloud key set created by this library
void flatten_bk(const LweBootstrappingKey* bk, int32_t* out);
void flatten_ksk(const LweKeySwitchKey* ks, int32_t* out);
bool flat_params_of(const TFheGateBootstrappingParameterSet* p, tfhe_flat_params* out);
// time-domain bk of an upstream-shaped FFT-domain key (values in LagrangeHalfCPolynomial.data)
int bk_from_lagrange_struct(const LweBootstrappingKeyFFT* bkfft, const tfhe_flat_params* p, int32_t* bk);

// device key bound to a cloud key set (created on first use)
tfhe_gpu_key* device_key_for(const TFheGateBootstrappingCloudKeySet* bk);
tfhe_gpu_key* device_key_for_bkfft(const TGswSampleFFT* bkfft);
tfhe_gpu_key* device_key_for_lwebk(const LweBootstrappingKeyFFT* bkfft);
tfhe_gpu_key* device_key_for_ksk(const LweKeySwitchKey* ks);
void register_keyset(const TFheGateBootstrappingCloudKeySet* bk);
void forget_device_keys(const TFheGateBootstrappingCloudKeySet* bk);
void key_params(const tfhe_gpu_key* key, tfhe_flat_params* out);
int32_t default_precision();

// ---- key-set containers shared by keygen and file import (tfhe_api.cpp, tfhe_io.cpp) ----
TFheGateBootstrappingCloudKeySet alloc_cloud_keyset(const TFheGateBootstrappingParameterSet* params);
LweKey* alloc_lwe_key(const LweParams* params);
TGswKey* alloc_tgsw_key(const TGswParams* params);
TFheGateBootstrappingParameterSet* alloc_params(const tfhe_flat_params& f, double lwe_alpha_min, double lwe_alpha_max,
                                                double tlwe_alpha_min, double tlwe_alpha_max);
// ownership records of objects read from files: a standalone cloud key set owns its parameter
// set; a secret key set read from a file owns its parameter set.  take_* removes the record.
void note_standalone_cloud(const TFheGateBootstrappingCloudKeySet* c, const TFheGateBootstrappingParameterSet* p);
bool take_standalone_cloud(const TFheGateBootstrappingCloudKeySet* c, const TFheGateBootstrappingParameterSet** p);
void note_owned_params(const TFheGateBootstrappingSecretKeySet* k, const TFheGateBootstrappingParameterSet* p);
const TFheGateBootstrappingParameterSet* take_owned_params(const TFheGateBootstrappingSecretKeySet* k);

// ---- runtime services shared with the circuit layer (all on the library's stream) ----
struct DevKeyView;
struct DevTwiddles;
std::recursive_mutex& lib_mutex();
int lib_ensure_init();                       // TFHE_OK or an error code (last error set)
hipStream_t lib_stream();
int lib_twiddles(int N, DevTwiddles** out);
int lib_hip_fail(const char* what, hipError_t e);
int lib_arg_fail(const char* what);
int key_check_prec(tfhe_gpu_key* key, int32_t prec);
DevKeyView key_view(const tfhe_gpu_key* key);
// creation number of a device key (unique in the process; never reused after the key is freed)
uint64_t key_id(const tfhe_gpu_key* key);
// gate pre-combination x = (0, C) + alpha a + beta b of the two-input gates
void gate_coefficients(int gate, int32_t* C, int32_t* alpha, int32_t* beta);
// per-launch timing under tfhe_gpu_profile_enable (a no-op otherwise)
struct LibProf {
    explicit LibProf(const char* name);
    ~LibProf();
    void* impl;
};

}  // namespace tfhe
