// icbc.h -- drop-in replacement for castano/icbc v1.05's public API, backed by
// the MI355X (gfx950) BC1 encoder in libicbc_amd.so.
//
// Same namespace, enums, defaults and signatures as the reference header's
// public section (icbc.h:10-52): every function the reference declares there
// (init, decode_bc1, decode_bc3, evaluate_bc1_error, evaluate_bc3_error,
// compress_bc1, compress_bc3) is here, so code that calls them compiles and
// links against libicbc_amd.so without change.  The reference's internal
// functions (everything inside its ICBC_IMPLEMENTATION section) are not part
// of this API.  Every call runs on the GPU.  The reference API has no error
// channel: after a device failure, or before a successful init, the float
// functions return NaN, outputs are left untouched, and icbc_gpu_last_error()
// (icbc_gpu.h) gives the reason.  For throughput, encode many blocks per call
// through icbc_gpu.h instead of block by block.
#ifndef ICBC_H
#define ICBC_H

namespace icbc {

    enum Decoder {
        Decoder_D3D10 = 0,
        Decoder_NVIDIA = 1,
        Decoder_AMD = 2,
        Decoder_Intel = 3
    };

    // Reference: icbc.h:19, 3813.  Builds the encoder tables for `decoder`
    // and uploads them to the current device.  Not thread-safe.
    void init(Decoder decoder = Decoder_D3D10);

    enum Quality {
        Quality_Level0,  // Box fit only.
        Quality_Level1,  // Box fit + least squares fit.
        Quality_Level2,  // Cluster fit 4, threshold = 24.
        Quality_Level3,  // Cluster fit 4, threshold = 32.
        Quality_Level4,  // Cluster fit 4, threshold = 48.
        Quality_Level5,  // Cluster fit 4, threshold = 64.
        Quality_Level6,  // Cluster fit 4, threshold = 96.
        Quality_Level7,  // Cluster fit 4, threshold = 128.
        Quality_Level8,  // Cluster fit 4+3, threshold = 256.
        Quality_Level9,  // Cluster fit 4+3, threshold = 256 + Refinement.

        Quality_Fast = Quality_Level1,
        Quality_Default = Quality_Level8,
        Quality_Max = Quality_Level9,
    };

    // Reference: icbc.h:38, 3820.
    void decode_bc1(const void * block, unsigned char rgba_block[16 * 4], Decoder decoder = Decoder_D3D10);

    // Reference: icbc.h:39, 3824 (2929-2946).  16-byte BC3 block: colours
    // from the four-colour palette of its BC1 part, every texel's alpha =
    // alpha0 (the reference's alpha indices are unimplemented).
    void decode_bc3(const void * block, unsigned char rgba_block[16 * 4], Decoder decoder = Decoder_D3D10);

    // Reference: icbc.h:43, 3837.
    float evaluate_bc1_error(const unsigned char rgba_block[16 * 4], const void * block, Decoder decoder = Decoder_D3D10);

    // Reference: icbc.h:44, 3841 (3048-3071), with the same decoding as
    // decode_bc3; the colour error is scaled by alpha / 255 when alpha_blend.
    float evaluate_bc3_error(const unsigned char rgba_block[16 * 4], const void * block, bool alpha_blend, Decoder decoder = Decoder_D3D10);

    // Reference: icbc.h:48, 3846.  Returns the weighted squared error of the
    // encoded block; FLT_MAX with `output` untouched for an invalid level.
    float compress_bc1(Quality level, const float * input_colors, const float * input_weights, const float color_weights[3], bool three_color_mode, bool three_color_black, void * output);

    // Reference: icbc.h:49, 3850.  The reference's BC3 encoder is a stub that
    // encodes nothing and returns 0; this keeps that behaviour.
    float compress_bc3(Quality level, const float * input_colors, const float * input_weights, const float color_weights[3], bool six_alpha_mode, void * output);
}

#endif // ICBC_H
