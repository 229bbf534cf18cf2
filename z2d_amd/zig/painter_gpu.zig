// painter_gpu.zig - z2d's painter entry points on the MI355X backend.
//
// Drop this file next to src/painter.zig and route `painter.fill` /
// `painter.stroke` (src/painter.zig:66-72, :214-220) and
// `SurfaceCompositor.run` (src/compositor.zig:302-309) here.  The signatures
// are the reference's; the argument unions are copied field by field into the
// `extern struct` descriptors of include/z2d_types.h (Zig unions have no
// guaranteed layout) and handed to libz2d_gpu.so through include/z2d_gpu.h.
//
// Build (where a Zig toolchain exists; none is in this repository's build):
//   exe.addIncludePath(b.path("include"));
//   exe.addLibraryPath(b.path("z2d_amd"));
//   exe.linkSystemLibrary("z2d_gpu");
//
// Status codes come back in the reference's error names; the device-side
// conditions the reference cannot raise keep names of their own (GpuError)
// instead of collapsing into OutOfMemory.

const std = @import("std");
const mem = std.mem;

const c = @cImport(@cInclude("z2d_gpu.h"));

const z2d = @import("z2d.zig");
const painter = z2d.painter;
const compositor = z2d.compositor;
const pixel = z2d.pixel;
const Surface = z2d.Surface;
const Pattern = z2d.Pattern;
const Gradient = z2d.Gradient;
const Dither = z2d.Dither;
const color = z2d.color;
const PathNode = @import("internal/path_nodes.zig").PathNode;

/// Conditions of the device path with no counterpart in the reference.
pub const GpuError = error{
    /// A HIP call or a kernel failed (Z2D_ERR_DEVICE).
    DeviceFault,
    /// The request can never fit the configured HBM budget, or the fixed
    /// scratch of the no-alloc entry point (Z2D_ERR_BUDGET_EXCEEDED).
    BudgetExceeded,
    /// No context could be created for this thread (Z2D_ERR_NO_CONTEXT).
    NoContext,
    /// A descriptor the marshaling below produced was rejected
    /// (Z2D_ERR_INVALID_ARGUMENT).
    InvalidArgument,
    /// A mode the device path does not implement, e.g. hairline strokes
    /// (Z2D_ERR_UNSUPPORTED).
    Unsupported,
};

pub const FillError = painter.FillError || GpuError;
pub const StrokeError = painter.StrokeError || GpuError;

// ---------------------------------------------------------------------------
// status codes (include/z2d_types.h, enum z2d_status)
// ---------------------------------------------------------------------------
fn gpuStatus(rc: i32) GpuError {
    return switch (rc) {
        c.Z2D_ERR_BUDGET_EXCEEDED => error.BudgetExceeded,
        c.Z2D_ERR_NO_CONTEXT => error.NoContext,
        c.Z2D_ERR_INVALID_ARGUMENT => error.InvalidArgument,
        c.Z2D_ERR_UNSUPPORTED => error.Unsupported,
        else => error.DeviceFault,
    };
}

fn fillStatus(rc: i32) FillError!void {
    return switch (rc) {
        c.Z2D_OK => {},
        c.Z2D_ERR_PATH_NOT_CLOSED => error.PathNotClosed,
        c.Z2D_ERR_PIXEL_SOURCE_NOT_PREMULTIPLIED => error.PixelSourceNotPreMultiplied,
        c.Z2D_ERR_INVALID_WIDTH => error.InvalidWidth,
        c.Z2D_ERR_INVALID_HEIGHT => error.InvalidHeight,
        c.Z2D_ERR_INVALID_STATE => error.InvalidState,
        c.Z2D_ERR_OUT_OF_MEMORY => error.OutOfMemory,
        else => gpuStatus(rc),
    };
}

fn strokeStatus(rc: i32) StrokeError!void {
    return switch (rc) {
        c.Z2D_OK => {},
        c.Z2D_ERR_PIXEL_SOURCE_NOT_PREMULTIPLIED => error.PixelSourceNotPreMultiplied,
        c.Z2D_ERR_INVALID_MATRIX => error.InvalidMatrix,
        c.Z2D_ERR_INVALID_WIDTH => error.InvalidWidth,
        c.Z2D_ERR_INVALID_HEIGHT => error.InvalidHeight,
        c.Z2D_ERR_INVALID_STATE => error.InvalidState,
        c.Z2D_ERR_OUT_OF_MEMORY => error.OutOfMemory,
        else => gpuStatus(rc),
    };
}

// ---------------------------------------------------------------------------
// one context per thread (it owns a HIP stream)
// ---------------------------------------------------------------------------
threadlocal var thread_ctx: ?*c.z2d_gpu_ctx = null;

fn threadContext() GpuError!*c.z2d_gpu_ctx {
    if (thread_ctx) |ctx| return ctx;
    var ctx: ?*c.z2d_gpu_ctx = null;
    const rc = c.z2d_gpu_create(0, &ctx);
    if (rc != c.Z2D_OK or ctx == null) return gpuStatus(if (rc == c.Z2D_OK) c.Z2D_ERR_NO_CONTEXT else rc);
    thread_ctx = ctx;
    return ctx.?;
}

/// Releases the calling thread's context (optional; contexts otherwise live
/// until the process ends).
pub fn deinitThread() void {
    if (thread_ctx) |ctx| c.z2d_gpu_destroy(ctx);
    thread_ctx = null;
}

// ---------------------------------------------------------------------------
// marshaling
// ---------------------------------------------------------------------------
fn surfaceDesc(surface: *Surface) c.z2d_surface_desc {
    return switch (surface.*) {
        inline else => |*s, tag| .{
            .format = @intFromEnum(@field(pixel.Format, @tagName(tag)["image_surface_".len..])),
            .width = s.width,
            .height = s.height,
            .reserved = 0,
            .data = @ptrCast(s.buf.ptr),
            .len = mem.sliceAsBytes(s.buf).len,
        },
    };
}

fn pixelDesc(px: pixel.Pixel) c.z2d_pixel {
    return switch (px) {
        inline .argb, .rgba => |p, tag| .{ .format = @intFromEnum(tag), .r = p.r, .g = p.g, .b = p.b, .a = p.a },
        inline .xrgb, .rgb => |p, tag| .{ .format = @intFromEnum(tag), .r = p.r, .g = p.g, .b = p.b, .a = 0 },
        inline .alpha8, .alpha4, .alpha2, .alpha1 => |p, tag| .{ .format = @intFromEnum(tag), .r = 0, .g = 0, .b = 0, .a = p.a },
    };
}

fn colorDesc(col: color.Color) c.z2d_color {
    return switch (col) {
        .linear_rgb => |v| .{ .space = c.Z2D_COLOR_LINEAR_RGB, .v = .{ v.r, v.g, v.b, v.a } },
        .srgb => |v| .{ .space = c.Z2D_COLOR_SRGB, .v = .{ v.r, v.g, v.b, v.a } },
        .hsl => |v| .{ .space = c.Z2D_COLOR_HSL, .v = .{ v.h, v.s, v.l, v.a } },
    };
}

fn colorArgsDesc(args: color.Color.InitArgs) c.z2d_color {
    return colorDesc(color.Color.init(args));
}

/// A gradient and its stops (which the descriptor points into).
const GradientDesc = struct {
    desc: c.z2d_gradient_desc,
    stops: []c.z2d_stop,

    fn init(alloc: mem.Allocator, g: *const Gradient) mem.Allocator.Error!GradientDesc {
        var d: c.z2d_gradient_desc = std.mem.zeroes(c.z2d_gradient_desc);
        var items: []const z2d.gradient.Stop = undefined;
        var method: color.InterpolationMethod = undefined;
        var m: z2d.Transformation = undefined;
        switch (g.*) {
            .linear => |l| {
                d.type = c.Z2D_GRADIENT_LINEAR;
                d.geom = .{ l.start.x, l.start.y, l.end.x, l.end.y, 0, 0 };
                items = l.stops.l.items;
                method = l.stops.interpolation_method;
                m = l.transformation;
            },
            .radial => |r| {
                d.type = c.Z2D_GRADIENT_RADIAL;
                d.geom = .{ r.inner.x, r.inner.y, r.inner_radius, r.outer.x, r.outer.y, r.outer_radius };
                items = r.stops.l.items;
                method = r.stops.interpolation_method;
                m = r.transformation;
            },
            .conic => |k| {
                d.type = c.Z2D_GRADIENT_CONIC;
                d.geom = .{ k.center.x, k.center.y, k.angle, 0, 0, 0 };
                items = k.stops.l.items;
                method = k.stops.interpolation_method;
                m = k.transformation;
            },
        }
        switch (method) {
            .linear_rgb => d.method = c.Z2D_INTERP_LINEAR_RGB,
            .srgb => d.method = c.Z2D_INTERP_SRGB,
            .hsl => |polar| {
                d.method = c.Z2D_INTERP_HSL;
                d.polar = @intFromEnum(polar);
            },
        }
        // the stored (already inverted) transformation, as the device reads it
        d.matrix = .{ m.ax, m.by, m.cx, m.dy, m.tx, m.ty };
        // stops in list order: addAssumeCapacity keeps them sorted, stable
        const stops = try alloc.alloc(c.z2d_stop, items.len);
        for (items, stops) |s, *o| o.* = .{ .offset = s.offset, .color = colorDesc(s.color) };
        d.n_stops = @intCast(stops.len);
        d.stops = stops.ptr;
        return .{ .desc = d, .stops = stops };
    }

    fn deinit(self: *GradientDesc, alloc: mem.Allocator) void {
        alloc.free(self.stops);
    }
};

/// A pattern descriptor and the gradient it may point to.
const PatternDesc = struct {
    desc: c.z2d_pattern_desc,
    grad: ?*GradientDesc = null,

    fn init(alloc: mem.Allocator, p: *const Pattern) mem.Allocator.Error!PatternDesc {
        var out: PatternDesc = .{ .desc = std.mem.zeroes(c.z2d_pattern_desc) };
        switch (p.*) {
            .opaque_pattern => |o| {
                out.desc.kind = c.Z2D_PATTERN_OPAQUE;
                out.desc.pixel = pixelDesc(o.pixel);
            },
            .gradient => |g| {
                out.grad = try gradientBox(alloc, g);
                out.desc.kind = c.Z2D_PATTERN_GRADIENT;
                out.desc.gradient = &out.grad.?.desc;
            },
            .dither => |d| {
                out.desc.kind = c.Z2D_PATTERN_DITHER;
                out.desc.dither.type = @intFromEnum(d.type);
                out.desc.dither.scale = d.scale;
                switch (d.source) {
                    .pixel => |px| {
                        out.desc.dither.source = c.Z2D_DITHER_SRC_PIXEL;
                        out.desc.dither.pixel = pixelDesc(px);
                    },
                    .color => |args| {
                        out.desc.dither.source = c.Z2D_DITHER_SRC_COLOR;
                        out.desc.dither.color = colorArgsDesc(args);
                    },
                    .gradient => |g| {
                        out.grad = try gradientBox(alloc, g);
                        out.desc.dither.source = c.Z2D_DITHER_SRC_GRADIENT;
                        out.desc.dither.gradient = &out.grad.?.desc;
                    },
                }
            },
        }
        return out;
    }

    fn gradientBox(alloc: mem.Allocator, g: *const Gradient) mem.Allocator.Error!*GradientDesc {
        const box = try alloc.create(GradientDesc);
        errdefer alloc.destroy(box);
        box.* = try GradientDesc.init(alloc, g);
        return box;
    }

    fn deinit(self: *PatternDesc, alloc: mem.Allocator) void {
        if (self.grad) |g| {
            g.deinit(alloc);
            alloc.destroy(g);
        }
    }
};

fn nodeArray(alloc: mem.Allocator, nodes: []const PathNode) mem.Allocator.Error![]c.z2d_path_node {
    const out = try alloc.alloc(c.z2d_path_node, nodes.len);
    for (nodes, out) |n, *o| o.* = switch (n) {
        .move_to => |m| .{ .tag = c.Z2D_NODE_MOVE_TO, .reserved = 0, .v = .{ m.point.x, m.point.y, 0, 0, 0, 0 } },
        .line_to => |l| .{ .tag = c.Z2D_NODE_LINE_TO, .reserved = 0, .v = .{ l.point.x, l.point.y, 0, 0, 0, 0 } },
        .curve_to => |q| .{ .tag = c.Z2D_NODE_CURVE_TO, .reserved = 0, .v = .{ q.p1.x, q.p1.y, q.p2.x, q.p2.y, q.p3.x, q.p3.y } },
        .close_path => .{ .tag = c.Z2D_NODE_CLOSE_PATH, .reserved = 0, .v = .{ 0, 0, 0, 0, 0, 0 } },
    };
    return out;
}

// ---------------------------------------------------------------------------
// entry points
// ---------------------------------------------------------------------------

/// painter.fill (src/painter.zig:66-143) on the device.  The host buffer is
/// uploaded, rendered and read back before return.
pub fn fill(
    alloc: mem.Allocator,
    surface: *Surface,
    pattern: *const Pattern,
    nodes: []const PathNode,
    opts: painter.FillOptions,
) FillError!void {
    var sdesc = surfaceDesc(surface);
    var pdesc = try PatternDesc.init(alloc, pattern);
    defer pdesc.deinit(alloc);
    const cnodes = try nodeArray(alloc, nodes);
    defer alloc.free(cnodes);
    const copts: c.z2d_fill_opts = .{
        .anti_aliasing_mode = @intFromEnum(opts.anti_aliasing_mode),
        .fill_rule = @intFromEnum(opts.fill_rule),
        .op = @intFromEnum(opts.operator),
        .precision = @intFromEnum(opts.precision),
        .tolerance = opts.tolerance,
    };
    // zero nodes is a no-op before any context is needed, as in the reference
    if (nodes.len == 0) return fillStatus(c.z2d_gpu_fill(null, &sdesc, &pdesc.desc, cnodes.ptr, 0, &copts));
    const ctx = try threadContext();
    return fillStatus(c.z2d_gpu_fill(ctx, &sdesc, &pdesc.desc, cnodes.ptr, cnodes.len, &copts));
}

/// painter.stroke (src/painter.zig:214-344): host tessellation in the
/// library, then the device raster.
pub fn stroke(
    alloc: mem.Allocator,
    surface: *Surface,
    pattern: *const Pattern,
    nodes: []const PathNode,
    opts: painter.StrokeOptions,
) StrokeError!void {
    var sdesc = surfaceDesc(surface);
    var pdesc = try PatternDesc.init(alloc, pattern);
    defer pdesc.deinit(alloc);
    const cnodes = try nodeArray(alloc, nodes);
    defer alloc.free(cnodes);
    const t = opts.transformation;
    const copts: c.z2d_stroke_opts = .{
        .anti_aliasing_mode = @intFromEnum(opts.anti_aliasing_mode),
        .line_cap_mode = @intFromEnum(opts.line_cap_mode),
        .line_join_mode = @intFromEnum(opts.line_join_mode),
        .op = @intFromEnum(opts.operator),
        .precision = @intFromEnum(opts.precision),
        .hairline = @intFromBool(opts.hairline),
        .dashes = opts.dashes.ptr,
        .n_dashes = opts.dashes.len,
        .dash_offset = opts.dash_offset,
        .line_width = opts.line_width,
        .miter_limit = opts.miter_limit,
        .tolerance = opts.tolerance,
        .transformation = .{ t.ax, t.by, t.cx, t.dy, t.tx, t.ty },
    };
    const ctx = if (nodes.len == 0) null else try threadContext();
    return strokeStatus(c.z2d_gpu_stroke(ctx, &sdesc, &pdesc.desc, cnodes.ptr, cnodes.len, &copts));
}

/// SurfaceCompositor.run (src/compositor.zig:302-440) for surface, pixel,
/// gradient and dither parameters.
pub fn compositeRun(
    alloc: mem.Allocator,
    dst: *Surface,
    dst_x: i32,
    dst_y: i32,
    operations: []const compositor.SurfaceCompositor.Operation,
    run_opts: compositor.SurfaceCompositor.RunOptions,
) (GpuError || mem.Allocator.Error)!void {
    var ddesc = surfaceDesc(dst);
    const ops = try alloc.alloc(c.z2d_comp_op, operations.len);
    defer alloc.free(ops);
    // surface parameters need their descriptors to outlive the call
    const sdescs = try alloc.alloc(c.z2d_surface_desc, 2 * operations.len);
    defer alloc.free(sdescs);
    var pdescs: std.ArrayListUnmanaged(PatternDesc) = .empty;
    defer {
        for (pdescs.items) |*p| p.deinit(alloc);
        pdescs.deinit(alloc);
    }
    for (operations, ops, 0..) |op, *o, i| {
        o.* = std.mem.zeroes(c.z2d_comp_op);
        o.op = @intFromEnum(op.operator);
        try paramDesc(alloc, op.dst, &o.dst, &sdescs[2 * i], &pdescs);
        try paramDesc(alloc, op.src, &o.src, &sdescs[2 * i + 1], &pdescs);
    }
    const ctx = try threadContext();
    const rc = c.z2d_gpu_composite(ctx, &ddesc, dst_x, dst_y, ops.ptr, ops.len, @intFromEnum(run_opts.precision), null);
    if (rc != c.Z2D_OK) return gpuStatus(rc);
}

fn paramDesc(
    alloc: mem.Allocator,
    param: ?compositor.SurfaceCompositor.Operation.Param,
    out: *c.z2d_comp_param,
    sdesc: *c.z2d_surface_desc,
    pdescs: *std.ArrayListUnmanaged(PatternDesc),
) mem.Allocator.Error!void {
    const p = param orelse {
        out.kind = c.Z2D_PARAM_NONE;
        return;
    };
    switch (p) {
        .pixel => |px| {
            out.kind = c.Z2D_PARAM_PIXEL;
            out.pixel = pixelDesc(px);
        },
        .surface => |s| {
            sdesc.* = surfaceDesc(s);
            out.kind = c.Z2D_PARAM_SURFACE;
            out.surface = sdesc;
        },
        .gradient => |g| {
            const pat: Pattern = .{ .gradient = g };
            try pdescs.append(alloc, try PatternDesc.init(alloc, &pat));
            out.kind = c.Z2D_PARAM_GRADIENT;
            out.gradient = pdescs.items[pdescs.items.len - 1].desc.gradient;
        },
        .dither => |d| {
            const pat: Pattern = .{ .dither = d };
            try pdescs.append(alloc, try PatternDesc.init(alloc, &pat));
            out.kind = c.Z2D_PARAM_DITHER;
            out.dither = pdescs.items[pdescs.items.len - 1].desc.dither;
        },
    }
}
