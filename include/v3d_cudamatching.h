// -*- C++ -*-
// Drop-in replacement of V3D_CUDA::SIFT_Matching (V3D/CUDA/v3d_cudamatching.h:15-67) on
// top of the gfx950 matcher's C ABI (v3dm.h, libv3dmatch.so).
//
// Same member functions, argument meaning and defaults as the reference header.  The
// differences a caller can observe:
//   * errors throw std::runtime_error (the reference prints and exit()s, cu:12-27);
//   * callers no longer call cublasInit()/cublasShutdown() (match_sequence.cpp:111, :411);
//   * the guided-matching functions are templates over the matrix and position types, so
//     V3D::Matrix3x3d (M[r][c]) and V3D::Vector2f (p[0], p[1]) are accepted without
//     including the V3D math headers here;
//   * the object owns its device memory and releases it in the destructor as well as in
//     deallocate().
#ifndef V3D_CUDA_MATCHING_H
#define V3D_CUDA_MATCHING_H

#include <algorithm>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "v3dm.h"

namespace V3D_CUDA
{

   struct SIFT_Matching
   {
         SIFT_Matching()
            : _warnOnRounding(true), _useReductionPath(false), _useGuidedMatching(false),
              _nMaxLeftFeatures(0), _nMaxRightFeatures(0), _nLeftFeatures(0), _nRightFeatures(0), _h(0)
         { }

         ~SIFT_Matching() { if (_h) v3dm_destroy(_h); }

         SIFT_Matching(SIFT_Matching const&) = delete;
         SIFT_Matching& operator=(SIFT_Matching const&) = delete;

         void enableRoundingWarning(bool flag = true)
         {
            _warnOnRounding = flag;
            if (_h) check(v3dm_set_flags(_h, _warnOnRounding, _useReductionPath));
         }

         void enableReductionPath(bool flag = true)
         {
            _useReductionPath = flag;
            if (_h) check(v3dm_set_flags(_h, _warnOnRounding, _useReductionPath));
         }

         void allocate(int nMaxLeftFeatures, int nMaxRightFeatures, bool useGuidedMatching = false)
         {
            handle();
            check(v3dm_allocate(_h, nMaxLeftFeatures, nMaxRightFeatures, useGuidedMatching ? 1 : 0));
            _useGuidedMatching = useGuidedMatching;
            _nMaxLeftFeatures = nMaxLeftFeatures - (nMaxLeftFeatures & 31);
            _nMaxRightFeatures = nMaxRightFeatures - (nMaxRightFeatures & 31);
         }

         void allocate(int nMaxFeatures, bool useGuidedMatching = false)
         {
            this->allocate(nMaxFeatures, nMaxFeatures, useGuidedMatching);
         }

         void deallocate()
         {
            if (_h) check(v3dm_deallocate(_h));
            _nLeftFeatures = _nRightFeatures = 0;
         }

         void setLeftFeatures(int nFeatures, float const * descriptors)
         {
            handle();
            check(v3dm_set_left(_h, nFeatures, descriptors));
            _nLeftFeatures = std::min(nFeatures - (nFeatures & 31), _nMaxLeftFeatures);
         }

         void setRightFeatures(int nFeatures, float const * descriptors)
         {
            handle();
            check(v3dm_set_right(_h, nFeatures, descriptors));
            _nRightFeatures = std::min(nFeatures - (nFeatures & 31), _nMaxRightFeatures);
         }

         // Ratio test is disabled if minRatio < 0.
         void findPutativeMatches(std::vector<std::pair<int, int> >& matches,
                                  float const minScore = 0.9f, float const minRatio = 0.8f)
         {
            handle();
            int const cap = std::max(1, std::min(_nLeftFeatures, _nRightFeatures));
            std::vector<int32_t> buf(2 * (size_t)cap);
            int n = 0;
            check(v3dm_find_putative_matches(_h, minScore, minRatio, buf.data(), cap, &n));
            fill(matches, buf, n);
         }

         template <typename Mat3, typename Vec2>
         void runGuidedMatching(Mat3 const& F, float const distanceThreshold,
                                Vec2 const * leftPositions, Vec2 const * rightPositions,
                                std::vector<std::pair<int, int> >& matches,
                                float const minScore = 0.9f, float const minRatio = 0.8f)
         {
            guided(0, F, distanceThreshold, leftPositions, rightPositions, matches, minScore, minRatio);
         }

         template <typename Mat3, typename Vec2>
         void runGuidedHomographyMatching(Mat3 const& H, float const distanceThreshold,
                                          Vec2 const * leftPositions, Vec2 const * rightPositions,
                                          std::vector<std::pair<int, int> >& matches,
                                          float const minScore = 0.9f, float const minRatio = 0.8f)
         {
            guided(1, H, distanceThreshold, leftPositions, rightPositions, matches, minScore, minRatio);
         }

         // Library extensions (no counterpart in the reference).
         void setPrecision(int mode) { handle(); check(v3dm_set_precision(_h, mode)); }
         v3dm_handle nativeHandle() { handle(); return _h; }

      protected:
         void handle()
         {
            if (_h) return;
            if (v3dm_create(&_h, 0) != V3DM_OK)
            {
               _h = 0;
               throw std::runtime_error(std::string("SIFT_Matching: ") + v3dm_last_error(0));
            }
            check(v3dm_set_flags(_h, _warnOnRounding, _useReductionPath));
         }

         void check(int rc)
         {
            if (rc != V3DM_OK) throw std::runtime_error(std::string("SIFT_Matching: ") + v3dm_last_error(_h));
         }

         static void fill(std::vector<std::pair<int, int> >& matches, std::vector<int32_t> const& buf, int n)
         {
            matches.clear();
            matches.reserve(n);
            for (int k = 0; k < n; ++k) matches.push_back(std::make_pair(buf[2 * k], buf[2 * k + 1]));
         }

         template <typename Mat3, typename Vec2>
         void guided(int kind, Mat3 const& M, float const distanceThreshold,
                     Vec2 const * leftPositions, Vec2 const * rightPositions,
                     std::vector<std::pair<int, int> >& matches, float const minScore, float const minRatio)
         {
            handle();
            double m[9];
            for (int r = 0; r < 3; ++r)
               for (int c = 0; c < 3; ++c) m[3 * r + c] = M[r][c];
            std::vector<float> l(2 * (size_t)_nLeftFeatures), rr(2 * (size_t)_nRightFeatures);
            for (int k = 0; k < _nLeftFeatures; ++k) { l[2 * k] = leftPositions[k][0]; l[2 * k + 1] = leftPositions[k][1]; }
            for (int k = 0; k < _nRightFeatures; ++k) { rr[2 * k] = rightPositions[k][0]; rr[2 * k + 1] = rightPositions[k][1]; }
            int const cap = std::max(1, std::min(_nLeftFeatures, _nRightFeatures));
            std::vector<int32_t> buf(2 * (size_t)cap);
            int n = 0;
            if (kind == 0)
               check(v3dm_run_guided_F(_h, m, distanceThreshold, l.data(), rr.data(), minScore, minRatio,
                                       buf.data(), cap, &n));
            else
               check(v3dm_run_guided_H(_h, m, distanceThreshold, l.data(), rr.data(), minScore, minRatio,
                                       buf.data(), cap, &n));
            fill(matches, buf, n);
         }

         bool   _warnOnRounding;
         bool   _useReductionPath;
         bool   _useGuidedMatching;
         int    _nMaxLeftFeatures, _nMaxRightFeatures, _nLeftFeatures, _nRightFeatures;
         v3dm_handle _h;
   }; // end struct SIFT_Matching

   // Drop-in replacement of VocabularyTreeTraversal<BranchFactor, Depth>
   // (V3D/CUDA/v3d_cudamatching.h:69-105, cu:795-869): traverseTree writes the node id of
   // every feature's leaf, as the CUDA kernel does.  The descent follows the CPU tree's
   // rule (first child of maximal sequential-FP32 score per level, v3d_voctree.h:63-99);
   // the CUDA kernel's own summation order is not reproduced.  Errors throw.
   template <int BranchFactor, int Depth>
   struct VocabularyTreeTraversal
   {
         VocabularyTreeTraversal() : _nMaxFeatures(0), _t(0) { }
         ~VocabularyTreeTraversal() { if (_t) v3dm_voctree_destroy(_t); }

         VocabularyTreeTraversal(VocabularyTreeTraversal const&) = delete;
         VocabularyTreeTraversal& operator=(VocabularyTreeTraversal const&) = delete;

         void allocate(int nMaxFeatures)
         {
            if ((nMaxFeatures % 32) != 0)  // cu:806-811
            {
               std::fprintf(stderr, "VocabularyTreeTraversal::allocate() warning: nMaxFeatures should be a multiple of 32\n");
               std::fprintf(stderr, "Rounding nFeatures down.\n");
               nMaxFeatures = nMaxFeatures - (nMaxFeatures & 31);
            }
            _nMaxFeatures = nMaxFeatures;
            if (!_t) check(v3dm_voctree_create(&_t, 0, BranchFactor, Depth));
         }

         void deallocate()
         {
            if (_t) v3dm_voctree_destroy(_t);
            _t = 0;
         }

         void setNodes(float const * nodes)
         {
            if (!_t) check(v3dm_voctree_create(&_t, 0, BranchFactor, Depth));
            check(v3dm_voctree_set_nodes(_t, nodes));
         }

         void traverseTree(int nFeatures, float const * features, int * leaveIds)
         {
            if (nFeatures > _nMaxFeatures)  // cu:843-848
            {
               std::fprintf(stderr, "VocabularyTreeTraversal::traverseTree() warning: nFeatures is larger than nMaxFeatures.\n");
               std::fprintf(stderr, "Reducing nFeatures.\n");
               nFeatures = _nMaxFeatures;
            }
            if (!_t) throw std::runtime_error("VocabularyTreeTraversal: setNodes() first");
            check(v3dm_voctree_traverse(_t, nFeatures, features, leaveIds, 1));
         }

         static inline int getNodeCount()
         {
            int res = 0, len = 1;
            for (int d = 0; d < Depth; ++d) { len *= BranchFactor; res += len; }
            return res;
         }

         static inline int getLeaveCount()
         {
            int res = 1;
            for (int d = 0; d < Depth; ++d) res *= BranchFactor;
            return res;
         }

      protected:
         static void check(int rc)
         {
            if (rc != V3DM_OK) throw std::runtime_error(std::string("v3dm: ") + v3dm_last_error(0));
         }

         int          _nMaxFeatures;
         v3dm_voctree _t;
   }; // end struct VocabularyTreeTraversal

} // end namespace V3D_CUDA

#endif
