// P/Invoke binding of libsvo (include/svo.h) and a ComputeManager-shaped wrapper over it.
//
// Drop-in for SvoTracer.Kernel.ComputeManager (ComputeManager.cs:10-367): the same public
// method names and argument meanings, with buffers in HIP device memory instead of OpenCL
// buffers.  Every call runs on the library's single in-order HIP stream, so the event
// values returned for the reference's CLEvent-returning methods only order calls that are
// already ordered; Wait(...) waits for all work enqueued so far.
//
// Build: add this file to SvoTracer.Kernel and place libsvo.so (Linux) next to the
// executable.  Struct layouts are the natural OpenCL layouts (wavefront.cl:1-84); the
// host serialisers that pack some of them differently (UsageSerializer, PruningSerializer,
// GraftingSerializer, TraceInputDataSerializer, UpdateInputDataSerializer) are replaced by
// the explicit layouts below.
using System;
using System.Collections.Generic;
using System.Runtime.InteropServices;

namespace SvoTracer.Kernel.Native
{
	[StructLayout(LayoutKind.Explicit, Size = 104)]
	public struct NativeTraceInput
	{
		[FieldOffset(0)] public float OriginX;
		[FieldOffset(4)] public float OriginY;
		[FieldOffset(8)] public float OriginZ;
		[FieldOffset(12)] public unsafe fixed float Facing[9];
		[FieldOffset(48)] public float FoVX;
		[FieldOffset(52)] public float FoVY;
		[FieldOffset(56)] public float DoFX;
		[FieldOffset(60)] public float DoFY;
		[FieldOffset(64)] public uint ScreenWidth;
		[FieldOffset(68)] public uint ScreenHeight;
		[FieldOffset(72)] public byte MaxOpacity;
		[FieldOffset(73)] public byte BaseDepth;
		[FieldOffset(74)] public ushort Tick;
		[FieldOffset(76)] public uint MaxChildRequestId;
		[FieldOffset(80)] public float FovMultiplier;
		[FieldOffset(84)] public float FovConstant;
		[FieldOffset(88)] public float WeightingMultiplier;
		[FieldOffset(92)] public float WeightingConstant;
		[FieldOffset(96)] public byte AmbientLightLevel;
	}

	[StructLayout(LayoutKind.Explicit, Size = 20)]
	public struct NativeUpdateInputData
	{
		[FieldOffset(0)] public byte BaseDepth;
		[FieldOffset(2)] public ushort Tick;
		[FieldOffset(4)] public uint MaxChildRequestId;
		[FieldOffset(8)] public uint MemorySize;
		[FieldOffset(12)] public uint Offset;
		[FieldOffset(16)] public uint GraftSize;
	}

	[StructLayout(LayoutKind.Explicit, Size = 16)]
	public struct NativeBlock
	{
		[FieldOffset(0)] public uint Child;
		[FieldOffset(4)] public ushort Chunk;
		[FieldOffset(6)] public short NormalPitch;
		[FieldOffset(8)] public short NormalYaw;
		[FieldOffset(10)] public byte ColourR;
		[FieldOffset(11)] public byte ColourG;
		[FieldOffset(12)] public byte ColourB;
		[FieldOffset(13)] public byte Opacity;
		[FieldOffset(14)] public byte Specularity;
		[FieldOffset(15)] public byte Gloss;
	}

	[StructLayout(LayoutKind.Explicit, Size = 8)]
	public struct NativeUsage
	{
		[FieldOffset(0)] public ushort Tick;
		[FieldOffset(4)] public uint Parent;
	}

	[StructLayout(LayoutKind.Explicit, Size = 20)]
	public struct NativePruning
	{
		[FieldOffset(0)] public byte Properties;
		[FieldOffset(1)] public byte Depth;
		[FieldOffset(4)] public uint Address;
		[FieldOffset(8)] public ushort Chunk;
		[FieldOffset(12)] public uint ColourAddress;
		[FieldOffset(16)] public uint ChildAddress;
	}

	[StructLayout(LayoutKind.Explicit, Size = 16)]
	public struct NativeGrafting
	{
		[FieldOffset(0)] public uint GraftDataAddress;
		[FieldOffset(4)] public uint GraftTotalSize;
		[FieldOffset(8)] public byte Depth;
		[FieldOffset(12)] public uint GraftAddress;
	}

	[StructLayout(LayoutKind.Explicit, Size = 40)]
	public struct NativeChildRequest
	{
		[FieldOffset(0)] public uint Address;
		[FieldOffset(4)] public ushort Tick;
		[FieldOffset(6)] public byte Depth;
		[FieldOffset(8)] public ulong LocationX;
		[FieldOffset(16)] public ulong LocationY;
		[FieldOffset(24)] public ulong LocationZ;
		[FieldOffset(32)] public byte TreeSize;
	}

	[StructLayout(LayoutKind.Sequential)]
	public struct NativeTreeStats
	{
		public uint Applied, Skipped, Invalid, Requests, Collected, Placed, Unlinks, Reserved;
	}

	[StructLayout(LayoutKind.Sequential)]
	public struct NativeFrameOptions
	{
		public uint Flags, MaxTreeDepth, WorkgroupSize, RefillLanes, StackLevels, WorkChunk, WavesPerSimd, Reserved;
	}

	// NativeFrameOptions.Flags (SVO_FRAME_* of svo.h)
	public static class FrameFlags
	{
		public const uint State = 1, Queues = 2, NoXcdSplit = 4, CameraTable = 8, RasterOrder = 16, KeepRequests = 32;
	}

	[StructLayout(LayoutKind.Sequential)]
	public struct NativeCompositeOptions
	{
		public uint MaxLayers, Flags, Reserved0, Reserved1;
	}

	[StructLayout(LayoutKind.Sequential)]
	public struct NativeAdamOptions
	{
		public float LearningRate, Beta1, Beta2, Epsilon;
		public uint Step;
	}

	[StructLayout(LayoutKind.Explicit, Size = 32)]
	public struct NativeStreamOptions
	{
		[FieldOffset(0)] public uint BaseDepth;
		[FieldOffset(4)] public uint MaxDepth;
		[FieldOffset(8)] public ushort Tick;
		[FieldOffset(12)] public uint MaxBricks;
		[FieldOffset(16)] public uint MaxChildRequestId;
		[FieldOffset(20)] public uint PruneAge;
		[FieldOffset(24)] public uint GcFreePercent;
		[FieldOffset(28)] public uint Flags;
	}

	[StructLayout(LayoutKind.Sequential)]
	public struct NativeStreamStats
	{
		public uint Requests, Unique, Bricks, Collected, GcRuns, Reclaimed, Moved, FreeGroups, PoolGroups, Reserved0, Reserved1, Reserved2;
	}

	/// <summary>svo_geometry: scene description of the host builder and the device terrain.</summary>
	[StructLayout(LayoutKind.Sequential)]
	public unsafe struct NativeGeometry
	{
		public uint Kind, Count;
		public float* Shapes;
		public ulong Seed;
		public uint Octaves, Resolution;
		public float Lacunarity, Gain, Amplitude, Frequency, HalfExtent, ZMin, ZMax;
	}

	/// <summary>Raw entry points of include/svo.h.</summary>
	public static unsafe class SvoNative
	{
		const string Lib = "svo";

		[DllImport(Lib)] public static extern int svo_device_count(out int count);
		[DllImport(Lib)] public static extern int svo_create(int deviceOrdinal, out IntPtr ctx);
		[DllImport(Lib)] public static extern int svo_destroy(IntPtr ctx);
		[DllImport(Lib)] public static extern IntPtr svo_last_error();
		[DllImport(Lib)] public static extern int svo_abi_version();

		[DllImport(Lib)] public static extern int svo_buffer_init(IntPtr ctx, int bufferName, nuint bytes, void* hostData, uint flags);
		[DllImport(Lib)] public static extern int svo_buffer_release(IntPtr ctx, int bufferName);
		[DllImport(Lib)] public static extern int svo_buffer_size(IntPtr ctx, int bufferName, out nuint bytes);
		[DllImport(Lib)] public static extern int svo_buffer_write(IntPtr ctx, int bufferName, nuint offset, void* src, nuint bytes);
		[DllImport(Lib)] public static extern int svo_buffer_read(IntPtr ctx, int bufferName, nuint offset, void* dst, nuint bytes);
		[DllImport(Lib)] public static extern int svo_buffer_fill(IntPtr ctx, int bufferName, void* pattern, nuint patternBytes, nuint offset, nuint bytes);
		[DllImport(Lib)] public static extern int svo_buffer_copy(IntPtr ctx, int dstName, nuint dstOffset, int srcName, nuint srcOffset, nuint bytes);
		[DllImport(Lib)] public static extern int svo_reset_id(IntPtr ctx, int bufferName, out uint count);
		[DllImport(Lib)] public static extern int svo_zero_id(IntPtr ctx, int bufferName);
		[DllImport(Lib)] public static extern int svo_swap(IntPtr ctx, int bufferA, int bufferB);
		[DllImport(Lib)] public static extern int svo_buffer_device_pointer(IntPtr ctx, int bufferName, out IntPtr devicePtr);

		[DllImport(Lib)] public static extern int svo_set_arg_buffer(IntPtr ctx, int kernelName, [MarshalAs(UnmanagedType.LPStr)] string param, int bufferName);
		[DllImport(Lib)] public static extern int svo_set_arg_bytes(IntPtr ctx, int kernelName, [MarshalAs(UnmanagedType.LPStr)] string param, void* value, nuint bytes);
		[DllImport(Lib)] public static extern int svo_trace_frame_static(IntPtr ctx, ref NativeTraceInput input, ref NativeFrameOptions options);
		[DllImport(Lib)] public static extern int svo_set_local_arg(IntPtr ctx, int kernelName, [MarshalAs(UnmanagedType.LPStr)] string param, int bufferName);
		[DllImport(Lib)] public static extern int svo_enqueue(IntPtr ctx, int kernelName, uint dims, nuint* globalSize);
		[DllImport(Lib)] public static extern int svo_flush(IntPtr ctx);
		[DllImport(Lib)] public static extern int svo_finish(IntPtr ctx);

		[DllImport(Lib)] public static extern int svo_trace_frame(IntPtr ctx, ref NativeTraceInput input, ref NativeFrameOptions options);
		[DllImport(Lib)] public static extern int svo_prune(IntPtr ctx, ref NativeUpdateInputData input, NativePruning* records, uint n,
			NativeBlock* surface, uint nSurface, ulong* locations, uint nLocations, out NativeTreeStats stats);
		[DllImport(Lib)] public static extern int svo_graft(IntPtr ctx, ref NativeUpdateInputData input, NativeGrafting* records, uint n,
			NativeBlock* graftBlocks, uint nGraftBlocks, ulong* locations, uint nLocations, out NativeTreeStats stats);
		[DllImport(Lib)] public static extern int svo_collect(IntPtr ctx, ref NativeUpdateInputData input, uint need, uint* groups, out uint count);
		[DllImport(Lib)] public static extern int svo_reclaim(IntPtr ctx, ref NativeUpdateInputData input, out uint reclaimed);
		[DllImport(Lib)] public static extern int svo_compact(IntPtr ctx, ref NativeUpdateInputData input, out uint moved);

		[DllImport(Lib)] public static extern int svo_trace_batch(IntPtr ctx, NativeTraceInput* inputs, float* apertures, uint nCameras,
			ref NativeFrameOptions options);
		[DllImport(Lib)] public static extern int svo_frame_graph_create(IntPtr ctx, NativeTraceInput* inputs, float* apertures, uint nCameras,
			ref NativeFrameOptions options, uint perCamera, out IntPtr graph);
		[DllImport(Lib)] public static extern int svo_frame_graph_launch(IntPtr ctx, IntPtr graph, NativeTraceInput* inputs);
		[DllImport(Lib)] public static extern int svo_frame_graph_destroy(IntPtr ctx, IntPtr graph);
		[DllImport(Lib)] public static extern int svo_autotune(IntPtr ctx, NativeTraceInput* inputs, float* apertures, uint nCameras, uint frames,
			ref NativeFrameOptions baseOptions, out NativeFrameOptions best, out float bestMs);
		[DllImport(Lib)] public static extern int svo_composite(IntPtr ctx, NativeTraceInput* views, float* apertures, uint nViews,
			ref NativeCompositeOptions options, out double loss);
		[DllImport(Lib)] public static extern int svo_adam_step(IntPtr ctx, ref NativeAdamOptions options);

		[DllImport(Lib)] public static extern int svo_terrain_create(IntPtr ctx, ref NativeGeometry geometry, out IntPtr terrain);
		[DllImport(Lib)] public static extern int svo_terrain_destroy(IntPtr terrain);
		[DllImport(Lib)] public static extern int svo_stream_seed(IntPtr ctx, IntPtr terrain, uint baseDepth);
		[DllImport(Lib)] public static extern int svo_terrain_build_tree(IntPtr ctx, IntPtr terrain, uint baseDepth, uint maxDepth,
			uint maxSize, int withUsages, ushort usageTick, out IntPtr tree, out uint blockCount);
		[DllImport(Lib)] public static extern int svo_tree_build_device(IntPtr ctx, ref NativeGeometry geometry, uint baseDepth,
			uint maxDepth, uint maxSize, int withUsages, ushort usageTick, out IntPtr tree, out uint blockCount);
		[DllImport(Lib)] public static extern int svo_stream_update(IntPtr ctx, IntPtr terrain, ref NativeStreamOptions options,
			out NativeStreamStats stats);
	}

	/// <summary>Stands in for CLEvent: the library's stream is in order, so events only mark calls.</summary>
	public readonly struct SvoEvent
	{
		public static readonly SvoEvent Done = new SvoEvent();
	}

	/// <summary>ComputeManager (ComputeManager.cs:10-367) over libsvo.</summary>
	public sealed unsafe class HipComputeManager : IDisposable
	{
		private IntPtr _ctx;

		public HipComputeManager(int deviceOrdinal = 0)
		{
			HandleResultCode(SvoNative.svo_create(deviceOrdinal, out _ctx), "svo_create");
		}

		public static void HandleResultCode(int resultCode, string method)
		{
			if (resultCode != 0)
				throw new Exception($"{method}: {resultCode} ({Marshal.PtrToStringAnsi(SvoNative.svo_last_error())})");
		}

		// ---- kernels (ComputeManager.cs:72-120) ----
		public SvoEvent Enqueue(KernelName kernelName, nuint[] workSize, SvoEvent[] waitEvents = null)
		{
			fixed (nuint* ws = workSize)
				HandleResultCode(SvoNative.svo_enqueue(_ctx, (int)kernelName, (uint)workSize.Length, ws), "EnqueueNDRangeKernel");
			return SvoEvent.Done;
		}

		public void SetArg<T>(KernelName kernelName, string paramName, T argument) where T : unmanaged
		{
			HandleResultCode(SvoNative.svo_set_arg_bytes(_ctx, (int)kernelName, paramName, &argument, (nuint)sizeof(T)),
				$"SetKernelArg:{kernelName}:{paramName}");
		}

		public void SetArg<T>(KernelName kernelName, string paramName, T[] argument) where T : unmanaged
		{
			fixed (T* p = argument)
				HandleResultCode(SvoNative.svo_set_arg_bytes(_ctx, (int)kernelName, paramName, p, (nuint)(argument.Length * sizeof(T))),
					$"SetKernelArg:{kernelName}:{paramName}");
		}

		public void SetArg(KernelName kernelName, string paramName, BufferName bufferName)
		{
			HandleResultCode(SvoNative.svo_set_arg_buffer(_ctx, (int)kernelName, paramName, (int)bufferName),
				$"SetKernelArg:{kernelName}:{paramName}");
		}

		public void SetLocalArg(KernelName kernelName, string paramName, BufferName bufferName)
		{
			HandleResultCode(SvoNative.svo_set_local_arg(_ctx, (int)kernelName, paramName, (int)bufferName),
				$"SetKernelArg:{kernelName}:{paramName}");
		}

		public void Wait(SvoEvent waitEvent) => HandleResultCode(SvoNative.svo_finish(_ctx), "WaitForEvents");
		public void Wait(SvoEvent[] waitEvents) { if (waitEvents != null) HandleResultCode(SvoNative.svo_finish(_ctx), "WaitForEvents"); }
		public void Flush() => HandleResultCode(SvoNative.svo_flush(_ctx), "Flush");

		// ---- buffers (ComputeManager.cs:124-275); sizes are bytes ----
		public void InitBuffer<T>(BufferName bufferName, T[] data) where T : unmanaged
		{
			fixed (T* p = data)
				HandleResultCode(SvoNative.svo_buffer_init(_ctx, (int)bufferName, (nuint)(data.Length * sizeof(T)), p, 0), $"CreateBuffer:{bufferName}");
		}

		public void InitReadBuffer<T>(BufferName bufferName, T[] data) where T : unmanaged
		{
			fixed (T* p = data)
				HandleResultCode(SvoNative.svo_buffer_init(_ctx, (int)bufferName, (nuint)(data.Length * sizeof(T)), p, 1), $"CreateBuffer:{bufferName}");
		}

		public void InitDeviceBuffer(BufferName bufferName, int arraySize, int elementSize)
		{
			HandleResultCode(SvoNative.svo_buffer_init(_ctx, (int)bufferName, (nuint)((long)arraySize * elementSize), null, 0), $"CreateBuffer:{bufferName}");
		}

		public void InitDeviceBuffer<T>(BufferName bufferName, int arraySize) where T : unmanaged
			=> InitDeviceBuffer(bufferName, arraySize, sizeof(T));

		public SvoEvent InitDeviceBuffer<T>(BufferName bufferName, int arraySize, T val) where T : unmanaged
		{
			InitDeviceBuffer(bufferName, arraySize, sizeof(T));
			HandleResultCode(SvoNative.svo_buffer_fill(_ctx, (int)bufferName, &val, (nuint)sizeof(T), 0, (nuint)((long)arraySize * sizeof(T))),
				$"EnqueueFillBuffer:{bufferName}");
			return SvoEvent.Done;
		}

		public void InitWriteBuffer<T>(BufferName bufferName, int arraySize) where T : unmanaged
		{
			HandleResultCode(SvoNative.svo_buffer_init(_ctx, (int)bufferName, (nuint)((long)arraySize * sizeof(T)), null, 2), $"CreateBuffer:{bufferName}");
		}

		public SvoEvent WriteBuffer<T>(BufferName bufferName, T[] data, SvoEvent[] waitEvents = null) where T : unmanaged
		{
			fixed (T* p = data)
				HandleResultCode(SvoNative.svo_buffer_write(_ctx, (int)bufferName, 0, p, (nuint)(data.Length * sizeof(T))), $"EnqueueWriteBuffer:{bufferName}");
			return SvoEvent.Done;
		}

		public void ReadBuffer<T>(BufferName bufferName, T[] array) where T : unmanaged
		{
			fixed (T* p = array)
				HandleResultCode(SvoNative.svo_buffer_read(_ctx, (int)bufferName, 0, p, (nuint)(array.Length * sizeof(T))), $"EnqueueReadBuffer:{bufferName}");
		}

		public (uint, SvoEvent) ResetIDBuffer(BufferName bufferName, SvoEvent[] initEvent)
		{
			HandleResultCode(SvoNative.svo_reset_id(_ctx, (int)bufferName, out uint count), $"EnqueueReadBuffer:{bufferName}");
			return (count, SvoEvent.Done);
		}

		public SvoEvent ZeroIDBuffer(BufferName bufferName, SvoEvent[] initEvent)
		{
			HandleResultCode(SvoNative.svo_zero_id(_ctx, (int)bufferName), $"EnqueueWriteBuffer:{bufferName}");
			return SvoEvent.Done;
		}

		public void Swap(BufferName buffer1, BufferName buffer2)
			=> HandleResultCode(SvoNative.svo_swap(_ctx, (int)buffer1, (int)buffer2), "Swap");

		public SvoEvent FillBuffer<T>(BufferName bufferName, T[] pattern, uint size) where T : unmanaged
		{
			fixed (T* p = pattern)
				HandleResultCode(SvoNative.svo_buffer_fill(_ctx, (int)bufferName, p, (nuint)(pattern.Length * sizeof(T)), 0, size), $"EnqueueFillBuffer:{bufferName}");
			return SvoEvent.Done;
		}

		public SvoEvent FillBuffer<T>(BufferName bufferName, T val) where T : unmanaged
		{
			HandleResultCode(SvoNative.svo_buffer_size(_ctx, (int)bufferName, out nuint bytes), $"EnqueueFillBuffer:{bufferName}");
			HandleResultCode(SvoNative.svo_buffer_fill(_ctx, (int)bufferName, &val, (nuint)sizeof(T), 0, bytes), $"EnqueueFillBuffer:{bufferName}");
			return SvoEvent.Done;
		}

		// ---- frame-level entry points (no ComputeManager counterpart) ----
		public void TraceFrame(NativeTraceInput input, NativeFrameOptions options = default)
			=> HandleResultCode(SvoNative.svo_trace_frame(_ctx, ref input, ref options), "TraceFrame");

		// TraceFrame with its scratch in the library's static device array (no allocation).
		public void TraceFrameStatic(NativeTraceInput input, NativeFrameOptions options = default)
			=> HandleResultCode(SvoNative.svo_trace_frame_static(_ctx, ref input, ref options), "TraceFrameStatic");

		public NativeTreeStats Prune(NativeUpdateInputData input, NativePruning[] records, NativeBlock[] surface, ulong[] locations)
		{
			fixed (NativePruning* r = records) fixed (NativeBlock* s = surface) fixed (ulong* l = locations)
			{
				HandleResultCode(SvoNative.svo_prune(_ctx, ref input, r, (uint)(records?.Length ?? 0), s, (uint)(surface?.Length ?? 0),
					l, (uint)((locations?.Length ?? 0) / 3), out NativeTreeStats st), "Prune");
				return st;
			}
		}

		public NativeTreeStats Graft(NativeUpdateInputData input, NativeGrafting[] records, NativeBlock[] graftBlocks, ulong[] locations)
		{
			fixed (NativeGrafting* r = records) fixed (NativeBlock* b = graftBlocks) fixed (ulong* l = locations)
			{
				HandleResultCode(SvoNative.svo_graft(_ctx, ref input, r, (uint)(records?.Length ?? 0), b, (uint)(graftBlocks?.Length ?? 0),
					l, (uint)((locations?.Length ?? 0) / 3), out NativeTreeStats st), "Graft");
				return st;
			}
		}

		public uint Reclaim(NativeUpdateInputData input)
		{
			HandleResultCode(SvoNative.svo_reclaim(_ctx, ref input, out uint n), "Reclaim");
			return n;
		}

		public uint Compact(NativeUpdateInputData input)
		{
			HandleResultCode(SvoNative.svo_compact(_ctx, ref input, out uint n), "Compact");
			return n;
		}

		// ---- batched cameras, captured frames, tuning ----
		public void TraceBatch(NativeTraceInput[] inputs, float[] apertures, NativeFrameOptions options = default)
		{
			fixed (NativeTraceInput* i = inputs) fixed (float* a = apertures)
				HandleResultCode(SvoNative.svo_trace_batch(_ctx, i, a, (uint)inputs.Length, ref options), "TraceBatch");
		}

		public IntPtr CaptureFrame(NativeTraceInput[] inputs, float[] apertures, NativeFrameOptions options = default, bool perCamera = false)
		{
			fixed (NativeTraceInput* i = inputs) fixed (float* a = apertures)
			{
				HandleResultCode(SvoNative.svo_frame_graph_create(_ctx, i, a, (uint)inputs.Length, ref options, perCamera ? 1u : 0u,
					out IntPtr graph), "CaptureFrame");
				return graph;
			}
		}

		public void LaunchFrame(IntPtr graph, NativeTraceInput[] inputs)
		{
			fixed (NativeTraceInput* i = inputs)
				HandleResultCode(SvoNative.svo_frame_graph_launch(_ctx, graph, i), "LaunchFrame");
		}

		public void ReleaseFrame(IntPtr graph) => HandleResultCode(SvoNative.svo_frame_graph_destroy(_ctx, graph), "ReleaseFrame");

		public (NativeFrameOptions, float) Autotune(NativeTraceInput[] inputs, float[] apertures, uint frames = 3,
			NativeFrameOptions baseOptions = default)
		{
			fixed (NativeTraceInput* i = inputs) fixed (float* a = apertures)
			{
				HandleResultCode(SvoNative.svo_autotune(_ctx, i, a, (uint)inputs.Length, frames, ref baseOptions,
					out NativeFrameOptions best, out float ms), "Autotune");
				return (best, ms);
			}
		}

		// ---- differentiable compositing ----
		public double Composite(NativeTraceInput[] views, float[] apertures, NativeCompositeOptions options)
		{
			fixed (NativeTraceInput* v = views) fixed (float* a = apertures)
			{
				HandleResultCode(SvoNative.svo_composite(_ctx, v, a, (uint)views.Length, ref options, out double loss), "Composite");
				return loss;
			}
		}

		public void AdamStep(NativeAdamOptions options) => HandleResultCode(SvoNative.svo_adam_step(_ctx, ref options), "AdamStep");

		// ---- streaming ----
		public IntPtr CreateTerrain(NativeGeometry geometry)
		{
			HandleResultCode(SvoNative.svo_terrain_create(_ctx, ref geometry, out IntPtr terrain), "CreateTerrain");
			return terrain;
		}

		public void DestroyTerrain(IntPtr terrain) => HandleResultCode(SvoNative.svo_terrain_destroy(terrain), "DestroyTerrain");

		public void StreamSeed(IntPtr terrain, uint baseDepth)
			=> HandleResultCode(SvoNative.svo_stream_seed(_ctx, terrain, baseDepth), "StreamSeed");

		// TreeBuilder.BuildTree of a device terrain, bound as this context's Bases/Blocks(/Usages);
		// the returned handle is a host copy of the tree (free it with svo_tree_free).
		// TreeBuilder.BuildTree of any builder geometry on the device (see BuildTreeOnDevice).
		public IntPtr BuildTreeDevice(NativeGeometry geometry, uint baseDepth, uint maxDepth, uint maxSize, bool withUsages,
			ushort usageTick, out uint blockCount)
		{
			HandleResultCode(SvoNative.svo_tree_build_device(_ctx, ref geometry, baseDepth, maxDepth, maxSize,
				withUsages ? 1 : 0, usageTick, out IntPtr tree, out blockCount), "BuildTreeDevice");
			return tree;
		}

		public IntPtr BuildTreeOnDevice(IntPtr terrain, uint baseDepth, uint maxDepth, uint maxSize, bool withUsages,
			ushort usageTick, out uint blockCount)
		{
			HandleResultCode(SvoNative.svo_terrain_build_tree(_ctx, terrain, baseDepth, maxDepth, maxSize,
				withUsages ? 1 : 0, usageTick, out IntPtr tree, out blockCount), "BuildTreeOnDevice");
			return tree;
		}

		public NativeStreamStats StreamUpdate(IntPtr terrain, NativeStreamOptions options)
		{
			HandleResultCode(SvoNative.svo_stream_update(_ctx, terrain, ref options, out NativeStreamStats stats), "StreamUpdate");
			return stats;
		}

		public void Dispose()
		{
			if (_ctx != IntPtr.Zero)
			{
				SvoNative.svo_destroy(_ctx);
				_ctx = IntPtr.Zero;
			}
		}
	}
}
